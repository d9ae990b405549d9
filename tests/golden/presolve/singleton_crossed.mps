NAME          SINGCROSS
ROWS
 N  OBJ
 G  S1
 L  R2
COLUMNS
    X         OBJ          1.0   S1           2.0
    X         R2           1.0
    Y         OBJ          1.0   R2           1.0
RHS
    RHS       S1          10.0   R2           9.0
BOUNDS
 UP BND       X            3.0
ENDATA
