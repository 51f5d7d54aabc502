NAME          BLENDING
ROWS
 L F1
 L F2
 N COST
COLUMNS
    P1        F1            0.3        F2          0.5
    P1      COST           -8.0
    P2        F1            0.7        F2          0.5
    P2      COST          -10.0
RHS
    RESOURCES F1          120.0        F2        210.0
ENDATA
