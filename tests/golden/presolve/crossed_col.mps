NAME          CROSSCOL
ROWS
 N  OBJ
 L  R1
COLUMNS
    X         OBJ          1.0   R1           1.0
    Y         OBJ          1.0   R1           1.0
RHS
    RHS       R1           5.0
BOUNDS
 LO BND       X            4.0
 UP BND       X            3.0
ENDATA
