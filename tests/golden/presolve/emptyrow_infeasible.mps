NAME          EMPTYROW
ROWS
 N  OBJ
 E  R0
 L  R1
COLUMNS
    X         OBJ          1.0   R1           1.0
    Y         OBJ          1.0   R1           1.0
RHS
    RHS       R0           5.0   R1           4.0
ENDATA
