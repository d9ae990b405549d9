NAME          DBLFILL
ROWS
 N  OBJ
 E  D1
 L  R2
 G  R3
 L  R4
COLUMNS
    X1        OBJ          3.0   D1           2.0
    X1        R2           1.0
    X2        OBJ          1.0   D1           3.0
    X2        R3           1.0   R4           1.0
    X5        OBJ         -1.0   R2           1.0
    X5        R3           2.0   R4           1.0
    X6        OBJ          0.5   R3           1.0
    X6        R4           4.0   R2           2.0
RHS
    RHS       D1           6.0   R2           4.0
    RHS       R3           1.0   R4          20.0
BOUNDS
 UP BND       X1          10.0
 UP BND       X2          10.0
 UP BND       X5           7.0
 UP BND       X6           7.0
ENDATA
