NAME          SENSE
OBJSENSE
    MAX
ROWS
 N  obj
 L  c1
COLUMNS
    x         obj          1.0   c1           1.0
RHS
    rhs       c1          10.0
ENDATA
