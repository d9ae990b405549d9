NAME          SINGLE
ROWS
 N  OBJ
 L  S1
 G  S2
 E  S3
 L  S4
 L  R5
 G  R6
COLUMNS
    X1        OBJ          1.0   S1          -2.0
    X1        R5           1.0   R6           1.0
    X2        OBJ         -1.0   S2           3.0
    X2        R5           1.0
    X3        OBJ          2.0   S3           4.0
    X3        R6           1.0   R5           1.0
    X4        OBJ          1.0   S4           1.0
    X4        R6           2.0   R5           1.0
RHS
    RHS       S1          -4.0   S2           3.0
    RHS       S3           8.0   S4           6.0
    RHS       R5          30.0   R6           1.0
RANGES
    RNG       S4           4.0
BOUNDS
 UP BND       X1          10.0
 UP BND       X2          10.0
 UP BND       X3          10.0
 MI BND       X4
 UP BND       X4          10.0
ENDATA
