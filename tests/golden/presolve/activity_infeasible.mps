NAME          ACTINF
ROWS
 N  OBJ
 L  R1
 G  R2
COLUMNS
    X         OBJ          1.0   R1           1.0
    X         R2           1.0
    Y         OBJ          1.0   R1           1.0
    Y         R2           1.0
RHS
    RHS       R1         -1.0   R2           1.0
BOUNDS
 UP BND       X            3.0
 UP BND       Y            3.0
ENDATA
