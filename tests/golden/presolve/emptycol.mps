NAME          EMPTYCOL
ROWS
 N  OBJ
 L  R1
 G  R2
 E  R3
COLUMNS
    P         OBJ          4.0
    Q         OBJ         -2.0
    S         OBJ          0.0
    X         OBJ          1.0   R1           1.0
    X         R2           1.0
    Y         OBJ          1.0   R1           1.0
    Y         R2           2.0   R3           1.0
    W         OBJ          1.0   R3           1.0
    W         R1           1.0
RHS
    RHS       R1          10.0   R2           2.0
    RHS       R3           3.0
BOUNDS
 LO BND       P            1.5
 UP BND       P            7.0
 UP BND       Q            3.0
 MI BND       S
 UP BND       S            2.5
 UP BND       X            9.0
 UP BND       Y            9.0
 UP BND       W            9.0
ENDATA
