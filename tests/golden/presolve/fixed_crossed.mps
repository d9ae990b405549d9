NAME          FIXCROSS
ROWS
 N  OBJ
 E  R1
 L  R2
COLUMNS
    X         OBJ          1.0   R1           1.0
    X         R2           1.0
    Y         OBJ          1.0   R1           1.0
    Y         R2           1.0
RHS
    RHS       R1           5.0   R2           2.0
BOUNDS
 FX BND       X            4.0
 FX BND       Y            3.0
ENDATA
