NAME          EMPTYUNB
ROWS
 N  OBJ
 L  R1
COLUMNS
    Q         OBJ         -2.0
    X         OBJ          1.0   R1           1.0
    Y         OBJ          1.0   R1           1.0
RHS
    RHS       R1           5.0
ENDATA
