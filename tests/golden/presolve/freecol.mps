NAME          FREECOL
ROWS
 N  OBJ
 E  DEF
 L  R2
 G  R3
COLUMNS
    Z         OBJ          2.0   DEF          1.0
    A         OBJ          1.0   DEF          1.0
    A         R2           1.0   R3           1.0
    B         OBJ          1.5   DEF          2.0
    B         R2           1.0
    C         OBJ         -0.5   R2           3.0
    C         R3           1.0
RHS
    RHS       DEF          3.0   R2          12.0
    RHS       R3           1.0
BOUNDS
 FR BND       Z
 UP BND       A            8.0
 UP BND       B            6.0
 UP BND       C            4.0
ENDATA
