NAME          EMPTYMAX
OBJSENSE
    MAX
ROWS
 N  OBJ
 L  R1
 L  R2
COLUMNS
    P         OBJ          4.0
    X         OBJ          1.0   R1           1.0
    X         R2           1.0
    Y         OBJ          2.0   R1           1.0
    Y         R2           3.0
RHS
    RHS       R1           5.0   R2           9.0
BOUNDS
 LO BND       P            1.0
 UP BND       P            6.0
 UP BND       X            9.0
 UP BND       Y            9.0
ENDATA
