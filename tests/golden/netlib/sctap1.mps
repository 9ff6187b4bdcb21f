NAME          SCTAP1
ROWS
 N  OBJZZZZZ
 G  NCZZ1ZZ1
 G  NCZZ2ZZ1
 G  NCZZ3ZZ1
 G  NCZZ4ZZ1
 G  NCZZ5ZZ1
 G  NCZZ6ZZ1
 G  ACZZ1ZZ1
 G  ACZZ2ZZ1
 G  ACZZ3ZZ1
 G  ACZZ4ZZ1
 G  ACZZ5ZZ1
 G  ACZZ6ZZ1
 G  ACZZ7ZZ1
 G  ACZZ8ZZ1
 G  ACZZ9ZZ1
 G  ACZ10ZZ1
 G  ACZ11ZZ1
 G  ACZ12ZZ1
 E  CCZZ1ZZ1
 E  CCZZ2ZZ1
 E  CCZZ3ZZ1
 E  CCZZ4ZZ1
 E  CCZZ5ZZ1
 E  CCZZ6ZZ1
 E  CCZZ7ZZ1
 E  CCZZ8ZZ1
 E  CCZZ9ZZ1
 E  CCZ10ZZ1
 E  CCZ11ZZ1
 E  CCZ12ZZ1
 G  NCZZ1ZZ2
 G  NCZZ2ZZ2
 G  NCZZ3ZZ2
 G  NCZZ4ZZ2
 G  NCZZ5ZZ2
 G  NCZZ6ZZ2
 G  ACZZ1ZZ2
 G  ACZZ2ZZ2
 G  ACZZ3ZZ2
 G  ACZZ4ZZ2
 G  ACZZ5ZZ2
 G  ACZZ6ZZ2
 G  ACZZ7ZZ2
 G  ACZZ8ZZ2
 G  ACZZ9ZZ2
 G  ACZ10ZZ2
 G  ACZ11ZZ2
 G  ACZ12ZZ2
 E  CCZZ1ZZ2
 E  CCZZ2ZZ2
 E  CCZZ3ZZ2
 E  CCZZ4ZZ2
 E  CCZZ5ZZ2
 E  CCZZ6ZZ2
 E  CCZZ7ZZ2
 E  CCZZ8ZZ2
 E  CCZZ9ZZ2
 E  CCZ10ZZ2
 E  CCZ11ZZ2
 E  CCZ12ZZ2
 G  NCZZ1ZZ3
 G  NCZZ2ZZ3
 G  NCZZ3ZZ3
 G  NCZZ4ZZ3
 G  NCZZ5ZZ3
 G  NCZZ6ZZ3
 G  ACZZ1ZZ3
 G  ACZZ2ZZ3
 G  ACZZ3ZZ3
 G  ACZZ4ZZ3
 G  ACZZ5ZZ3
 G  ACZZ6ZZ3
 G  ACZZ7ZZ3
 G  ACZZ8ZZ3
 G  ACZZ9ZZ3
 G  ACZ10ZZ3
 G  ACZ11ZZ3
 G  ACZ12ZZ3
 E  CCZZ1ZZ3
 E  CCZZ2ZZ3
 E  CCZZ3ZZ3
 E  CCZZ4ZZ3
 E  CCZZ5ZZ3
 E  CCZZ6ZZ3
 E  CCZZ7ZZ3
 E  CCZZ8ZZ3
 E  CCZZ9ZZ3
 E  CCZ10ZZ3
 E  CCZ11ZZ3
 E  CCZ12ZZ3
 G  NCZZ1ZZ4
 G  NCZZ2ZZ4
 G  NCZZ3ZZ4
 G  NCZZ4ZZ4
 G  NCZZ5ZZ4
 G  NCZZ6ZZ4
 G  ACZZ1ZZ4
 G  ACZZ2ZZ4
 G  ACZZ3ZZ4
 G  ACZZ4ZZ4
 G  ACZZ5ZZ4
 G  ACZZ6ZZ4
 G  ACZZ7ZZ4
 G  ACZZ8ZZ4
 G  ACZZ9ZZ4
 G  ACZ10ZZ4
 G  ACZ11ZZ4
 G  ACZ12ZZ4
 E  CCZZ1ZZ4
 E  CCZZ2ZZ4
 E  CCZZ3ZZ4
 E  CCZZ4ZZ4
 E  CCZZ5ZZ4
 E  CCZZ6ZZ4
 E  CCZZ7ZZ4
 E  CCZZ8ZZ4
 E  CCZZ9ZZ4
 E  CCZ10ZZ4
 E  CCZ11ZZ4
 E  CCZ12ZZ4
 G  NCZZ1ZZ5
 G  NCZZ2ZZ5
 G  NCZZ3ZZ5
 G  NCZZ4ZZ5
 G  NCZZ5ZZ5
 G  NCZZ6ZZ5
 G  ACZZ1ZZ5
 G  ACZZ2ZZ5
 G  ACZZ3ZZ5
 G  ACZZ4ZZ5
 G  ACZZ5ZZ5
 G  ACZZ6ZZ5
 G  ACZZ7ZZ5
 G  ACZZ8ZZ5
 G  ACZZ9ZZ5
 G  ACZ10ZZ5
 G  ACZ11ZZ5
 G  ACZ12ZZ5
 E  CCZZ1ZZ5
 E  CCZZ2ZZ5
 E  CCZZ3ZZ5
 E  CCZZ4ZZ5
 E  CCZZ5ZZ5
 E  CCZZ6ZZ5
 E  CCZZ7ZZ5
 E  CCZZ8ZZ5
 E  CCZZ9ZZ5
 E  CCZ10ZZ5
 E  CCZ11ZZ5
 E  CCZ12ZZ5
 G  NCZZ1ZZ6
 G  NCZZ2ZZ6
 G  NCZZ3ZZ6
 G  NCZZ4ZZ6
 G  NCZZ5ZZ6
 G  NCZZ6ZZ6
 G  ACZZ1ZZ6
 G  ACZZ2ZZ6
 G  ACZZ3ZZ6
 G  ACZZ4ZZ6
 G  ACZZ5ZZ6
 G  ACZZ6ZZ6
 G  ACZZ7ZZ6
 G  ACZZ8ZZ6
 G  ACZZ9ZZ6
 G  ACZ10ZZ6
 G  ACZ11ZZ6
 G  ACZ12ZZ6
 E  CCZZ1ZZ6
 E  CCZZ2ZZ6
 E  CCZZ3ZZ6
 E  CCZZ4ZZ6
 E  CCZZ5ZZ6
 E  CCZZ6ZZ6
 E  CCZZ7ZZ6
 E  CCZZ8ZZ6
 E  CCZZ9ZZ6
 E  CCZ10ZZ6
 E  CCZ11ZZ6
 E  CCZ12ZZ6
 G  NCZZ1ZZ7
 G  NCZZ2ZZ7
 G  NCZZ3ZZ7
 G  NCZZ4ZZ7
 G  NCZZ5ZZ7
 G  NCZZ6ZZ7
 G  ACZZ1ZZ7
 G  ACZZ2ZZ7
 G  ACZZ3ZZ7
 G  ACZZ4ZZ7
 G  ACZZ5ZZ7
 G  ACZZ6ZZ7
 G  ACZZ7ZZ7
 G  ACZZ8ZZ7
 G  ACZZ9ZZ7
 G  ACZ10ZZ7
 G  ACZ11ZZ7
 G  ACZ12ZZ7
 E  CCZZ1ZZ7
 E  CCZZ2ZZ7
 E  CCZZ3ZZ7
 E  CCZZ4ZZ7
 E  CCZZ5ZZ7
 E  CCZZ6ZZ7
 E  CCZZ7ZZ7
 E  CCZZ8ZZ7
 E  CCZZ9ZZ7
 E  CCZ10ZZ7
 E  CCZ11ZZ7
 E  CCZ12ZZ7
 G  NCZZ1ZZ8
 G  NCZZ2ZZ8
 G  NCZZ3ZZ8
 G  NCZZ4ZZ8
 G  NCZZ5ZZ8
 G  NCZZ6ZZ8
 G  ACZZ1ZZ8
 G  ACZZ2ZZ8
 G  ACZZ3ZZ8
 G  ACZZ4ZZ8
 G  ACZZ5ZZ8
 G  ACZZ6ZZ8
 G  ACZZ7ZZ8
 G  ACZZ8ZZ8
 G  ACZZ9ZZ8
 G  ACZ10ZZ8
 G  ACZ11ZZ8
 G  ACZ12ZZ8
 E  CCZZ1ZZ8
 E  CCZZ2ZZ8
 E  CCZZ3ZZ8
 E  CCZZ4ZZ8
 E  CCZZ5ZZ8
 E  CCZZ6ZZ8
 E  CCZZ7ZZ8
 E  CCZZ8ZZ8
 E  CCZZ9ZZ8
 E  CCZ10ZZ8
 E  CCZ11ZZ8
 E  CCZ12ZZ8
 G  NCZZ1ZZ9
 G  NCZZ2ZZ9
 G  NCZZ3ZZ9
 G  NCZZ4ZZ9
 G  NCZZ5ZZ9
 G  NCZZ6ZZ9
 G  ACZZ1ZZ9
 G  ACZZ2ZZ9
 G  ACZZ3ZZ9
 G  ACZZ4ZZ9
 G  ACZZ5ZZ9
 G  ACZZ6ZZ9
 G  ACZZ7ZZ9
 G  ACZZ8ZZ9
 G  ACZZ9ZZ9
 G  ACZ10ZZ9
 G  ACZ11ZZ9
 G  ACZ12ZZ9
 E  CCZZ1ZZ9
 E  CCZZ2ZZ9
 E  CCZZ3ZZ9
 E  CCZZ4ZZ9
 E  CCZZ5ZZ9
 E  CCZZ6ZZ9
 E  CCZZ7ZZ9
 E  CCZZ8ZZ9
 E  CCZZ9ZZ9
 E  CCZ10ZZ9
 E  CCZ11ZZ9
 E  CCZ12ZZ9
 G  NCZZ1Z10
 G  NCZZ2Z10
 G  NCZZ3Z10
 G  NCZZ4Z10
 G  NCZZ5Z10
 G  NCZZ6Z10
 G  ACZZ1Z10
 G  ACZZ2Z10
 G  ACZZ3Z10
 G  ACZZ4Z10
 G  ACZZ5Z10
 G  ACZZ6Z10
 G  ACZZ7Z10
 G  ACZZ8Z10
 G  ACZZ9Z10
 G  ACZ10Z10
 G  ACZ11Z10
 G  ACZ12Z10
 E  CCZZ1Z10
 E  CCZZ2Z10
 E  CCZZ3Z10
 E  CCZZ4Z10
 E  CCZZ5Z10
 E  CCZZ6Z10
 E  CCZZ7Z10
 E  CCZZ8Z10
 E  CCZZ9Z10
 E  CCZ10Z10
 E  CCZ11Z10
 E  CCZ12Z10
COLUMNS
    Z1ZZ1ZZ1  CCZZ1ZZ1            1.
    Z2ZZ1ZZ1  OBJZZZZZ           10.   NCZZ1ZZ1           10.
    Z2ZZ1ZZ1  ACZZ1ZZ1           10.   CCZZ1ZZ1            1.
    Z2ZZ1ZZ1  NCZZ2ZZ2          -10.
    Z3ZZ1ZZ1  OBJZZZZZ           20.   NCZZ1ZZ1           20.
    Z3ZZ1ZZ1  ACZZ1ZZ1           20.   CCZZ1ZZ1            1.
    Z3ZZ1ZZ1  NCZZ1ZZ2           -5.   ACZZ1ZZ2           -5.
    Z3ZZ1ZZ1  NCZZ2ZZ2          -15.
    Z4ZZ1ZZ1  OBJZZZZZ            1.   NCZZ1ZZ1            1.
    Z4ZZ1ZZ1  ACZZ1ZZ1            1.   NCZZ1ZZ2           -1.
    Z4ZZ1ZZ1  ACZZ1ZZ2           -1.
    Z1ZZ2ZZ1  CCZZ2ZZ1            1.
    Z2ZZ2ZZ1  OBJZZZZZ           15.   NCZZ2ZZ1           15.
    Z2ZZ2ZZ1  ACZZ2ZZ1           15.   CCZZ2ZZ1            1.
    Z2ZZ2ZZ1  NCZZ3ZZ2          -15.
    Z3ZZ2ZZ1  OBJZZZZZ           25.   NCZZ2ZZ1           25.
    Z3ZZ2ZZ1  ACZZ2ZZ1           25.   CCZZ2ZZ1            1.
    Z3ZZ2ZZ1  NCZZ2ZZ2           -7.   ACZZ2ZZ2           -7.
    Z3ZZ2ZZ1  NCZZ3ZZ2          -18.
    Z4ZZ2ZZ1  OBJZZZZZ            1.   NCZZ2ZZ1            1.
    Z4ZZ2ZZ1  ACZZ2ZZ1            1.   NCZZ2ZZ2           -1.
    Z4ZZ2ZZ1  ACZZ2ZZ2           -1.
    Z1ZZ3ZZ1  CCZZ3ZZ1            1.
    Z2ZZ3ZZ1  OBJZZZZZ           10.   NCZZ2ZZ1           10.
    Z2ZZ3ZZ1  ACZZ3ZZ1           10.   CCZZ3ZZ1            1.
    Z2ZZ3ZZ1  NCZZ1ZZ2          -10.
    Z3ZZ3ZZ1  OBJZZZZZ           20.   NCZZ2ZZ1           20.
    Z3ZZ3ZZ1  ACZZ3ZZ1           20.   CCZZ3ZZ1            1.
    Z3ZZ3ZZ1  NCZZ2ZZ2           -5.   ACZZ3ZZ2           -5.
    Z3ZZ3ZZ1  NCZZ1ZZ2          -15.
    Z4ZZ3ZZ1  OBJZZZZZ            1.   NCZZ2ZZ1            1.
    Z4ZZ3ZZ1  ACZZ3ZZ1            1.   NCZZ2ZZ2           -1.
    Z4ZZ3ZZ1  ACZZ3ZZ2           -1.
    Z1ZZ4ZZ1  CCZZ4ZZ1            1.
    Z2ZZ4ZZ1  OBJZZZZZ           15.   NCZZ1ZZ1           15.
    Z2ZZ4ZZ1  ACZZ4ZZ1           15.   CCZZ4ZZ1            1.
    Z2ZZ4ZZ1  NCZZ4ZZ2          -15.
    Z3ZZ4ZZ1  OBJZZZZZ           25.   NCZZ1ZZ1           25.
    Z3ZZ4ZZ1  ACZZ4ZZ1           25.   CCZZ4ZZ1            1.
    Z3ZZ4ZZ1  NCZZ1ZZ2           -7.   ACZZ4ZZ2           -7.
    Z3ZZ4ZZ1  NCZZ4ZZ2          -18.
    Z4ZZ4ZZ1  OBJZZZZZ            1.   NCZZ1ZZ1            1.
    Z4ZZ4ZZ1  ACZZ4ZZ1            1.   NCZZ1ZZ2           -1.
    Z4ZZ4ZZ1  ACZZ4ZZ2           -1.
    Z1ZZ5ZZ1  CCZZ5ZZ1            1.
    Z2ZZ5ZZ1  OBJZZZZZ           15.   NCZZ2ZZ1           15.
    Z2ZZ5ZZ1  ACZZ5ZZ1           15.   CCZZ5ZZ1            1.
    Z2ZZ5ZZ1  NCZZ4ZZ2          -15.
    Z3ZZ5ZZ1  OBJZZZZZ           25.   NCZZ2ZZ1           25.
    Z3ZZ5ZZ1  ACZZ5ZZ1           25.   CCZZ5ZZ1            1.
    Z3ZZ5ZZ1  NCZZ2ZZ2           -7.   ACZZ5ZZ2           -7.
    Z3ZZ5ZZ1  NCZZ4ZZ2          -18.
    Z4ZZ5ZZ1  OBJZZZZZ            1.   NCZZ2ZZ1            1.
    Z4ZZ5ZZ1  ACZZ5ZZ1            1.   NCZZ2ZZ2           -1.
    Z4ZZ5ZZ1  ACZZ5ZZ2           -1.
    Z1ZZ6ZZ1  CCZZ6ZZ1            1.
    Z2ZZ6ZZ1  OBJZZZZZ           10.   NCZZ2ZZ1           10.
    Z2ZZ6ZZ1  ACZZ6ZZ1           10.   CCZZ6ZZ1            1.
    Z2ZZ6ZZ1  NCZZ5ZZ2          -10.
    Z3ZZ6ZZ1  OBJZZZZZ           20.   NCZZ2ZZ1           20.
    Z3ZZ6ZZ1  ACZZ6ZZ1           20.   CCZZ6ZZ1            1.
    Z3ZZ6ZZ1  NCZZ2ZZ2           -5.   ACZZ6ZZ2           -5.
    Z3ZZ6ZZ1  NCZZ5ZZ2          -15.
    Z4ZZ6ZZ1  OBJZZZZZ            1.   NCZZ2ZZ1            1.
    Z4ZZ6ZZ1  ACZZ6ZZ1            1.   NCZZ2ZZ2           -1.
    Z4ZZ6ZZ1  ACZZ6ZZ2           -1.
    Z1ZZ7ZZ1  CCZZ7ZZ1            1.
    Z2ZZ7ZZ1  OBJZZZZZ           10.   NCZZ3ZZ1           10.
    Z2ZZ7ZZ1  ACZZ7ZZ1           10.   CCZZ7ZZ1            1.
    Z2ZZ7ZZ1  NCZZ5ZZ2          -10.
    Z3ZZ7ZZ1  OBJZZZZZ           20.   NCZZ3ZZ1           20.
    Z3ZZ7ZZ1  ACZZ7ZZ1           20.   CCZZ7ZZ1            1.
    Z3ZZ7ZZ1  NCZZ3ZZ2           -5.   ACZZ7ZZ2           -5.
    Z3ZZ7ZZ1  NCZZ5ZZ2          -15.
    Z4ZZ7ZZ1  OBJZZZZZ            1.   NCZZ3ZZ1            1.
    Z4ZZ7ZZ1  ACZZ7ZZ1            1.   NCZZ3ZZ2           -1.
    Z4ZZ7ZZ1  ACZZ7ZZ2           -1.
    Z1ZZ8ZZ1  CCZZ8ZZ1            1.
    Z2ZZ8ZZ1  OBJZZZZZ           30.   NCZZ4ZZ1           30.
    Z2ZZ8ZZ1  ACZZ8ZZ1           30.   CCZZ8ZZ1            1.
    Z2ZZ8ZZ1  NCZZ5ZZ2          -30.
    Z3ZZ8ZZ1  OBJZZZZZ           40.   NCZZ4ZZ1           40.
    Z3ZZ8ZZ1  ACZZ8ZZ1           40.   CCZZ8ZZ1            1.
    Z3ZZ8ZZ1  NCZZ4ZZ2           -7.   ACZZ8ZZ2           -7.
    Z3ZZ8ZZ1  NCZZ5ZZ2          -33.
    Z4ZZ8ZZ1  OBJZZZZZ            1.   NCZZ4ZZ1            1.
    Z4ZZ8ZZ1  ACZZ8ZZ1            1.   NCZZ4ZZ2           -1.
    Z4ZZ8ZZ1  ACZZ8ZZ2           -1.
    Z1ZZ9ZZ1  CCZZ9ZZ1            1.
    Z2ZZ9ZZ1  OBJZZZZZ           30.   NCZZ5ZZ1           30.
    Z2ZZ9ZZ1  ACZZ9ZZ1           30.   CCZZ9ZZ1            1.
    Z2ZZ9ZZ1  NCZZ4ZZ2          -30.
    Z3ZZ9ZZ1  OBJZZZZZ           40.   NCZZ5ZZ1           40.
    Z3ZZ9ZZ1  ACZZ9ZZ1           40.   CCZZ9ZZ1            1.
    Z3ZZ9ZZ1  NCZZ5ZZ2           -7.   ACZZ9ZZ2           -7.
    Z3ZZ9ZZ1  NCZZ4ZZ2          -33.
    Z4ZZ9ZZ1  OBJZZZZZ            1.   NCZZ5ZZ1            1.
    Z4ZZ9ZZ1  ACZZ9ZZ1            1.   NCZZ5ZZ2           -1.
    Z4ZZ9ZZ1  ACZZ9ZZ2           -1.
    Z1Z10ZZ1  CCZ10ZZ1            1.
    Z2Z10ZZ1  OBJZZZZZ           10.   NCZZ4ZZ1           10.
    Z2Z10ZZ1  ACZ10ZZ1           10.   CCZ10ZZ1            1.
    Z2Z10ZZ1  NCZZ6ZZ2          -10.
    Z3Z10ZZ1  OBJZZZZZ           20.   NCZZ4ZZ1           20.
    Z3Z10ZZ1  ACZ10ZZ1           20.   CCZ10ZZ1            1.
    Z3Z10ZZ1  NCZZ4ZZ2           -5.   ACZ10ZZ2           -5.
    Z3Z10ZZ1  NCZZ6ZZ2          -15.
    Z4Z10ZZ1  OBJZZZZZ            1.   NCZZ4ZZ1            1.
    Z4Z10ZZ1  ACZ10ZZ1            1.   NCZZ4ZZ2           -1.
    Z4Z10ZZ1  ACZ10ZZ2           -1.
    Z1Z11ZZ1  CCZ11ZZ1            1.
    Z2Z11ZZ1  OBJZZZZZ           15.   NCZZ6ZZ1           15.
    Z2Z11ZZ1  ACZ11ZZ1           15.   CCZ11ZZ1            1.
    Z3Z11ZZ1  OBJZZZZZ           25.   NCZZ6ZZ1           25.
    Z3Z11ZZ1  ACZ11ZZ1           25.   CCZ11ZZ1            1.
    Z3Z11ZZ1  NCZZ6ZZ2           -7.   ACZ11ZZ2           -7.
    Z4Z11ZZ1  OBJZZZZZ            1.   NCZZ6ZZ1            1.
    Z4Z11ZZ1  ACZ11ZZ1            1.   NCZZ6ZZ2           -1.
    Z4Z11ZZ1  ACZ11ZZ2           -1.
    Z1Z12ZZ1  CCZ12ZZ1            1.
    Z2Z12ZZ1  OBJZZZZZ           50.   NCZZ5ZZ1           50.
    Z2Z12ZZ1  ACZ12ZZ1           50.   CCZ12ZZ1            1.
    Z3Z12ZZ1  OBJZZZZZ           80.   NCZZ5ZZ1           80.
    Z3Z12ZZ1  ACZ12ZZ1           80.   CCZ12ZZ1            1.
    Z3Z12ZZ1  NCZZ5ZZ2          -15.   ACZ12ZZ2          -15.
    Z4Z12ZZ1  OBJZZZZZ            1.   NCZZ5ZZ1            1.
    Z4Z12ZZ1  ACZ12ZZ1            1.   NCZZ5ZZ2           -1.
    Z4Z12ZZ1  ACZ12ZZ2           -1.
    Z1ZZ1ZZ2  CCZZ1ZZ2            1.
    Z2ZZ1ZZ2  OBJZZZZZ           10.   NCZZ1ZZ2           10.
    Z2ZZ1ZZ2  ACZZ1ZZ2           10.   CCZZ1ZZ2            1.
    Z2ZZ1ZZ2  NCZZ2ZZ3          -10.
    Z3ZZ1ZZ2  OBJZZZZZ           20.   NCZZ1ZZ2           20.
    Z3ZZ1ZZ2  ACZZ1ZZ2           20.   CCZZ1ZZ2            1.
    Z3ZZ1ZZ2  NCZZ1ZZ3           -5.   ACZZ1ZZ3           -5.
    Z3ZZ1ZZ2  NCZZ2ZZ3          -15.
    Z4ZZ1ZZ2  OBJZZZZZ            1.   NCZZ1ZZ2            1.
    Z4ZZ1ZZ2  ACZZ1ZZ2            1.   NCZZ1ZZ3           -1.
    Z4ZZ1ZZ2  ACZZ1ZZ3           -1.
    Z1ZZ2ZZ2  CCZZ2ZZ2            1.
    Z2ZZ2ZZ2  OBJZZZZZ           15.   NCZZ2ZZ2           15.
    Z2ZZ2ZZ2  ACZZ2ZZ2           15.   CCZZ2ZZ2            1.
    Z2ZZ2ZZ2  NCZZ3ZZ3          -15.
    Z3ZZ2ZZ2  OBJZZZZZ           25.   NCZZ2ZZ2           25.
    Z3ZZ2ZZ2  ACZZ2ZZ2           25.   CCZZ2ZZ2            1.
    Z3ZZ2ZZ2  NCZZ2ZZ3           -7.   ACZZ2ZZ3           -7.
    Z3ZZ2ZZ2  NCZZ3ZZ3          -18.
    Z4ZZ2ZZ2  OBJZZZZZ            1.   NCZZ2ZZ2            1.
    Z4ZZ2ZZ2  ACZZ2ZZ2            1.   NCZZ2ZZ3           -1.
    Z4ZZ2ZZ2  ACZZ2ZZ3           -1.
    Z1ZZ3ZZ2  CCZZ3ZZ2            1.
    Z2ZZ3ZZ2  OBJZZZZZ           10.   NCZZ2ZZ2           10.
    Z2ZZ3ZZ2  ACZZ3ZZ2           10.   CCZZ3ZZ2            1.
    Z2ZZ3ZZ2  NCZZ1ZZ3          -10.
    Z3ZZ3ZZ2  OBJZZZZZ           20.   NCZZ2ZZ2           20.
    Z3ZZ3ZZ2  ACZZ3ZZ2           20.   CCZZ3ZZ2            1.
    Z3ZZ3ZZ2  NCZZ2ZZ3           -5.   ACZZ3ZZ3           -5.
    Z3ZZ3ZZ2  NCZZ1ZZ3          -15.
    Z4ZZ3ZZ2  OBJZZZZZ            1.   NCZZ2ZZ2            1.
    Z4ZZ3ZZ2  ACZZ3ZZ2            1.   NCZZ2ZZ3           -1.
    Z4ZZ3ZZ2  ACZZ3ZZ3           -1.
    Z1ZZ4ZZ2  CCZZ4ZZ2            1.
    Z2ZZ4ZZ2  OBJZZZZZ           15.   NCZZ1ZZ2           15.
    Z2ZZ4ZZ2  ACZZ4ZZ2           15.   CCZZ4ZZ2            1.
    Z2ZZ4ZZ2  NCZZ4ZZ3          -15.
    Z3ZZ4ZZ2  OBJZZZZZ           25.   NCZZ1ZZ2           25.
    Z3ZZ4ZZ2  ACZZ4ZZ2           25.   CCZZ4ZZ2            1.
    Z3ZZ4ZZ2  NCZZ1ZZ3           -7.   ACZZ4ZZ3           -7.
    Z3ZZ4ZZ2  NCZZ4ZZ3          -18.
    Z4ZZ4ZZ2  OBJZZZZZ            1.   NCZZ1ZZ2            1.
    Z4ZZ4ZZ2  ACZZ4ZZ2            1.   NCZZ1ZZ3           -1.
    Z4ZZ4ZZ2  ACZZ4ZZ3           -1.
    Z1ZZ5ZZ2  CCZZ5ZZ2            1.
    Z2ZZ5ZZ2  OBJZZZZZ           15.   NCZZ2ZZ2           15.
    Z2ZZ5ZZ2  ACZZ5ZZ2           15.   CCZZ5ZZ2            1.
    Z2ZZ5ZZ2  NCZZ4ZZ3          -15.
    Z3ZZ5ZZ2  OBJZZZZZ           25.   NCZZ2ZZ2           25.
    Z3ZZ5ZZ2  ACZZ5ZZ2           25.   CCZZ5ZZ2            1.
    Z3ZZ5ZZ2  NCZZ2ZZ3           -7.   ACZZ5ZZ3           -7.
    Z3ZZ5ZZ2  NCZZ4ZZ3          -18.
    Z4ZZ5ZZ2  OBJZZZZZ            1.   NCZZ2ZZ2            1.
    Z4ZZ5ZZ2  ACZZ5ZZ2            1.   NCZZ2ZZ3           -1.
    Z4ZZ5ZZ2  ACZZ5ZZ3           -1.
    Z1ZZ6ZZ2  CCZZ6ZZ2            1.
    Z2ZZ6ZZ2  OBJZZZZZ           10.   NCZZ2ZZ2           10.
    Z2ZZ6ZZ2  ACZZ6ZZ2           10.   CCZZ6ZZ2            1.
    Z2ZZ6ZZ2  NCZZ5ZZ3          -10.
    Z3ZZ6ZZ2  OBJZZZZZ           20.   NCZZ2ZZ2           20.
    Z3ZZ6ZZ2  ACZZ6ZZ2           20.   CCZZ6ZZ2            1.
    Z3ZZ6ZZ2  NCZZ2ZZ3           -5.   ACZZ6ZZ3           -5.
    Z3ZZ6ZZ2  NCZZ5ZZ3          -15.
    Z4ZZ6ZZ2  OBJZZZZZ            1.   NCZZ2ZZ2            1.
    Z4ZZ6ZZ2  ACZZ6ZZ2            1.   NCZZ2ZZ3           -1.
    Z4ZZ6ZZ2  ACZZ6ZZ3           -1.
    Z1ZZ7ZZ2  CCZZ7ZZ2            1.
    Z2ZZ7ZZ2  OBJZZZZZ           10.   NCZZ3ZZ2           10.
    Z2ZZ7ZZ2  ACZZ7ZZ2           10.   CCZZ7ZZ2            1.
    Z2ZZ7ZZ2  NCZZ5ZZ3          -10.
    Z3ZZ7ZZ2  OBJZZZZZ           20.   NCZZ3ZZ2           20.
    Z3ZZ7ZZ2  ACZZ7ZZ2           20.   CCZZ7ZZ2            1.
    Z3ZZ7ZZ2  NCZZ3ZZ3           -5.   ACZZ7ZZ3           -5.
    Z3ZZ7ZZ2  NCZZ5ZZ3          -15.
    Z4ZZ7ZZ2  OBJZZZZZ            1.   NCZZ3ZZ2            1.
    Z4ZZ7ZZ2  ACZZ7ZZ2            1.   NCZZ3ZZ3           -1.
    Z4ZZ7ZZ2  ACZZ7ZZ3           -1.
    Z1ZZ8ZZ2  CCZZ8ZZ2            1.
    Z2ZZ8ZZ2  OBJZZZZZ           30.   NCZZ4ZZ2           30.
    Z2ZZ8ZZ2  ACZZ8ZZ2           30.   CCZZ8ZZ2            1.
    Z2ZZ8ZZ2  NCZZ5ZZ3          -30.
    Z3ZZ8ZZ2  OBJZZZZZ           40.   NCZZ4ZZ2           40.
    Z3ZZ8ZZ2  ACZZ8ZZ2           40.   CCZZ8ZZ2            1.
    Z3ZZ8ZZ2  NCZZ4ZZ3           -7.   ACZZ8ZZ3           -7.
    Z3ZZ8ZZ2  NCZZ5ZZ3          -33.
    Z4ZZ8ZZ2  OBJZZZZZ            1.   NCZZ4ZZ2            1.
    Z4ZZ8ZZ2  ACZZ8ZZ2            1.   NCZZ4ZZ3           -1.
    Z4ZZ8ZZ2  ACZZ8ZZ3           -1.
    Z1ZZ9ZZ2  CCZZ9ZZ2            1.
    Z2ZZ9ZZ2  OBJZZZZZ           30.   NCZZ5ZZ2           30.
    Z2ZZ9ZZ2  ACZZ9ZZ2           30.   CCZZ9ZZ2            1.
    Z2ZZ9ZZ2  NCZZ4ZZ3          -30.
    Z3ZZ9ZZ2  OBJZZZZZ           40.   NCZZ5ZZ2           40.
    Z3ZZ9ZZ2  ACZZ9ZZ2           40.   CCZZ9ZZ2            1.
    Z3ZZ9ZZ2  NCZZ5ZZ3           -7.   ACZZ9ZZ3           -7.
    Z3ZZ9ZZ2  NCZZ4ZZ3          -33.
    Z4ZZ9ZZ2  OBJZZZZZ            1.   NCZZ5ZZ2            1.
    Z4ZZ9ZZ2  ACZZ9ZZ2            1.   NCZZ5ZZ3           -1.
    Z4ZZ9ZZ2  ACZZ9ZZ3           -1.
    Z1Z10ZZ2  CCZ10ZZ2            1.
    Z2Z10ZZ2  OBJZZZZZ           10.   NCZZ4ZZ2           10.
    Z2Z10ZZ2  ACZ10ZZ2           10.   CCZ10ZZ2            1.
    Z2Z10ZZ2  NCZZ6ZZ3          -10.
    Z3Z10ZZ2  OBJZZZZZ           20.   NCZZ4ZZ2           20.
    Z3Z10ZZ2  ACZ10ZZ2           20.   CCZ10ZZ2            1.
    Z3Z10ZZ2  NCZZ4ZZ3           -5.   ACZ10ZZ3           -5.
    Z3Z10ZZ2  NCZZ6ZZ3          -15.
    Z4Z10ZZ2  OBJZZZZZ            1.   NCZZ4ZZ2            1.
    Z4Z10ZZ2  ACZ10ZZ2            1.   NCZZ4ZZ3           -1.
    Z4Z10ZZ2  ACZ10ZZ3           -1.
    Z1Z11ZZ2  CCZ11ZZ2            1.
    Z2Z11ZZ2  OBJZZZZZ           15.   NCZZ6ZZ2           15.
    Z2Z11ZZ2  ACZ11ZZ2           15.   CCZ11ZZ2            1.
    Z3Z11ZZ2  OBJZZZZZ           25.   NCZZ6ZZ2           25.
    Z3Z11ZZ2  ACZ11ZZ2           25.   CCZ11ZZ2            1.
    Z3Z11ZZ2  NCZZ6ZZ3           -7.   ACZ11ZZ3           -7.
    Z4Z11ZZ2  OBJZZZZZ            1.   NCZZ6ZZ2            1.
    Z4Z11ZZ2  ACZ11ZZ2            1.   NCZZ6ZZ3           -1.
    Z4Z11ZZ2  ACZ11ZZ3           -1.
    Z1Z12ZZ2  CCZ12ZZ2            1.
    Z2Z12ZZ2  OBJZZZZZ           50.   NCZZ5ZZ2           50.
    Z2Z12ZZ2  ACZ12ZZ2           50.   CCZ12ZZ2            1.
    Z3Z12ZZ2  OBJZZZZZ           80.   NCZZ5ZZ2           80.
    Z3Z12ZZ2  ACZ12ZZ2           80.   CCZ12ZZ2            1.
    Z3Z12ZZ2  NCZZ5ZZ3          -15.   ACZ12ZZ3          -15.
    Z4Z12ZZ2  OBJZZZZZ            1.   NCZZ5ZZ2            1.
    Z4Z12ZZ2  ACZ12ZZ2            1.   NCZZ5ZZ3           -1.
    Z4Z12ZZ2  ACZ12ZZ3           -1.
    Z1ZZ1ZZ3  CCZZ1ZZ3            1.
    Z2ZZ1ZZ3  OBJZZZZZ           10.   NCZZ1ZZ3           10.
    Z2ZZ1ZZ3  ACZZ1ZZ3           10.   CCZZ1ZZ3            1.
    Z2ZZ1ZZ3  NCZZ2ZZ4          -10.
    Z3ZZ1ZZ3  OBJZZZZZ           20.   NCZZ1ZZ3           20.
    Z3ZZ1ZZ3  ACZZ1ZZ3           20.   CCZZ1ZZ3            1.
    Z3ZZ1ZZ3  NCZZ1ZZ4           -5.   ACZZ1ZZ4           -5.
    Z3ZZ1ZZ3  NCZZ2ZZ4          -15.
    Z4ZZ1ZZ3  OBJZZZZZ            1.   NCZZ1ZZ3            1.
    Z4ZZ1ZZ3  ACZZ1ZZ3            1.   NCZZ1ZZ4           -1.
    Z4ZZ1ZZ3  ACZZ1ZZ4           -1.
    Z1ZZ2ZZ3  CCZZ2ZZ3            1.
    Z2ZZ2ZZ3  OBJZZZZZ           15.   NCZZ2ZZ3           15.
    Z2ZZ2ZZ3  ACZZ2ZZ3           15.   CCZZ2ZZ3            1.
    Z2ZZ2ZZ3  NCZZ3ZZ4          -15.
    Z3ZZ2ZZ3  OBJZZZZZ           25.   NCZZ2ZZ3           25.
    Z3ZZ2ZZ3  ACZZ2ZZ3           25.   CCZZ2ZZ3            1.
    Z3ZZ2ZZ3  NCZZ2ZZ4           -7.   ACZZ2ZZ4           -7.
    Z3ZZ2ZZ3  NCZZ3ZZ4          -18.
    Z4ZZ2ZZ3  OBJZZZZZ            1.   NCZZ2ZZ3            1.
    Z4ZZ2ZZ3  ACZZ2ZZ3            1.   NCZZ2ZZ4           -1.
    Z4ZZ2ZZ3  ACZZ2ZZ4           -1.
    Z1ZZ3ZZ3  CCZZ3ZZ3            1.
    Z2ZZ3ZZ3  OBJZZZZZ           10.   NCZZ2ZZ3           10.
    Z2ZZ3ZZ3  ACZZ3ZZ3           10.   CCZZ3ZZ3            1.
    Z2ZZ3ZZ3  NCZZ1ZZ4          -10.
    Z3ZZ3ZZ3  OBJZZZZZ           20.   NCZZ2ZZ3           20.
    Z3ZZ3ZZ3  ACZZ3ZZ3           20.   CCZZ3ZZ3            1.
    Z3ZZ3ZZ3  NCZZ2ZZ4           -5.   ACZZ3ZZ4           -5.
    Z3ZZ3ZZ3  NCZZ1ZZ4          -15.
    Z4ZZ3ZZ3  OBJZZZZZ            1.   NCZZ2ZZ3            1.
    Z4ZZ3ZZ3  ACZZ3ZZ3            1.   NCZZ2ZZ4           -1.
    Z4ZZ3ZZ3  ACZZ3ZZ4           -1.
    Z1ZZ4ZZ3  CCZZ4ZZ3            1.
    Z2ZZ4ZZ3  OBJZZZZZ           15.   NCZZ1ZZ3           15.
    Z2ZZ4ZZ3  ACZZ4ZZ3           15.   CCZZ4ZZ3            1.
    Z2ZZ4ZZ3  NCZZ4ZZ4          -15.
    Z3ZZ4ZZ3  OBJZZZZZ           25.   NCZZ1ZZ3           25.
    Z3ZZ4ZZ3  ACZZ4ZZ3           25.   CCZZ4ZZ3            1.
    Z3ZZ4ZZ3  NCZZ1ZZ4           -7.   ACZZ4ZZ4           -7.
    Z3ZZ4ZZ3  NCZZ4ZZ4          -18.
    Z4ZZ4ZZ3  OBJZZZZZ            1.   NCZZ1ZZ3            1.
    Z4ZZ4ZZ3  ACZZ4ZZ3            1.   NCZZ1ZZ4           -1.
    Z4ZZ4ZZ3  ACZZ4ZZ4           -1.
    Z1ZZ5ZZ3  CCZZ5ZZ3            1.
    Z2ZZ5ZZ3  OBJZZZZZ           15.   NCZZ2ZZ3           15.
    Z2ZZ5ZZ3  ACZZ5ZZ3           15.   CCZZ5ZZ3            1.
    Z2ZZ5ZZ3  NCZZ4ZZ4          -15.
    Z3ZZ5ZZ3  OBJZZZZZ           25.   NCZZ2ZZ3           25.
    Z3ZZ5ZZ3  ACZZ5ZZ3           25.   CCZZ5ZZ3            1.
    Z3ZZ5ZZ3  NCZZ2ZZ4           -7.   ACZZ5ZZ4           -7.
    Z3ZZ5ZZ3  NCZZ4ZZ4          -18.
    Z4ZZ5ZZ3  OBJZZZZZ            1.   NCZZ2ZZ3            1.
    Z4ZZ5ZZ3  ACZZ5ZZ3            1.   NCZZ2ZZ4           -1.
    Z4ZZ5ZZ3  ACZZ5ZZ4           -1.
    Z1ZZ6ZZ3  CCZZ6ZZ3            1.
    Z2ZZ6ZZ3  OBJZZZZZ           10.   NCZZ2ZZ3           10.
    Z2ZZ6ZZ3  ACZZ6ZZ3           10.   CCZZ6ZZ3            1.
    Z2ZZ6ZZ3  NCZZ5ZZ4          -10.
    Z3ZZ6ZZ3  OBJZZZZZ           20.   NCZZ2ZZ3           20.
    Z3ZZ6ZZ3  ACZZ6ZZ3           20.   CCZZ6ZZ3            1.
    Z3ZZ6ZZ3  NCZZ2ZZ4           -5.   ACZZ6ZZ4           -5.
    Z3ZZ6ZZ3  NCZZ5ZZ4          -15.
    Z4ZZ6ZZ3  OBJZZZZZ            1.   NCZZ2ZZ3            1.
    Z4ZZ6ZZ3  ACZZ6ZZ3            1.   NCZZ2ZZ4           -1.
    Z4ZZ6ZZ3  ACZZ6ZZ4           -1.
    Z1ZZ7ZZ3  CCZZ7ZZ3            1.
    Z2ZZ7ZZ3  OBJZZZZZ           10.   NCZZ3ZZ3           10.
    Z2ZZ7ZZ3  ACZZ7ZZ3           10.   CCZZ7ZZ3            1.
    Z2ZZ7ZZ3  NCZZ5ZZ4          -10.
    Z3ZZ7ZZ3  OBJZZZZZ           20.   NCZZ3ZZ3           20.
    Z3ZZ7ZZ3  ACZZ7ZZ3           20.   CCZZ7ZZ3            1.
    Z3ZZ7ZZ3  NCZZ3ZZ4           -5.   ACZZ7ZZ4           -5.
    Z3ZZ7ZZ3  NCZZ5ZZ4          -15.
    Z4ZZ7ZZ3  OBJZZZZZ            1.   NCZZ3ZZ3            1.
    Z4ZZ7ZZ3  ACZZ7ZZ3            1.   NCZZ3ZZ4           -1.
    Z4ZZ7ZZ3  ACZZ7ZZ4           -1.
    Z1ZZ8ZZ3  CCZZ8ZZ3            1.
    Z2ZZ8ZZ3  OBJZZZZZ           30.   NCZZ4ZZ3           30.
    Z2ZZ8ZZ3  ACZZ8ZZ3           30.   CCZZ8ZZ3            1.
    Z2ZZ8ZZ3  NCZZ5ZZ4          -30.
    Z3ZZ8ZZ3  OBJZZZZZ           40.   NCZZ4ZZ3           40.
    Z3ZZ8ZZ3  ACZZ8ZZ3           40.   CCZZ8ZZ3            1.
    Z3ZZ8ZZ3  NCZZ4ZZ4           -7.   ACZZ8ZZ4           -7.
    Z3ZZ8ZZ3  NCZZ5ZZ4          -33.
    Z4ZZ8ZZ3  OBJZZZZZ            1.   NCZZ4ZZ3            1.
    Z4ZZ8ZZ3  ACZZ8ZZ3            1.   NCZZ4ZZ4           -1.
    Z4ZZ8ZZ3  ACZZ8ZZ4           -1.
    Z1ZZ9ZZ3  CCZZ9ZZ3            1.
    Z2ZZ9ZZ3  OBJZZZZZ           30.   NCZZ5ZZ3           30.
    Z2ZZ9ZZ3  ACZZ9ZZ3           30.   CCZZ9ZZ3            1.
    Z2ZZ9ZZ3  NCZZ4ZZ4          -30.
    Z3ZZ9ZZ3  OBJZZZZZ           40.   NCZZ5ZZ3           40.
    Z3ZZ9ZZ3  ACZZ9ZZ3           40.   CCZZ9ZZ3            1.
    Z3ZZ9ZZ3  NCZZ5ZZ4           -7.   ACZZ9ZZ4           -7.
    Z3ZZ9ZZ3  NCZZ4ZZ4          -33.
    Z4ZZ9ZZ3  OBJZZZZZ            1.   NCZZ5ZZ3            1.
    Z4ZZ9ZZ3  ACZZ9ZZ3            1.   NCZZ5ZZ4           -1.
    Z4ZZ9ZZ3  ACZZ9ZZ4           -1.
    Z1Z10ZZ3  CCZ10ZZ3            1.
    Z2Z10ZZ3  OBJZZZZZ           10.   NCZZ4ZZ3           10.
    Z2Z10ZZ3  ACZ10ZZ3           10.   CCZ10ZZ3            1.
    Z2Z10ZZ3  NCZZ6ZZ4          -10.
    Z3Z10ZZ3  OBJZZZZZ           20.   NCZZ4ZZ3           20.
    Z3Z10ZZ3  ACZ10ZZ3           20.   CCZ10ZZ3            1.
    Z3Z10ZZ3  NCZZ4ZZ4           -5.   ACZ10ZZ4           -5.
    Z3Z10ZZ3  NCZZ6ZZ4          -15.
    Z4Z10ZZ3  OBJZZZZZ            1.   NCZZ4ZZ3            1.
    Z4Z10ZZ3  ACZ10ZZ3            1.   NCZZ4ZZ4           -1.
    Z4Z10ZZ3  ACZ10ZZ4           -1.
    Z1Z11ZZ3  CCZ11ZZ3            1.
    Z2Z11ZZ3  OBJZZZZZ           15.   NCZZ6ZZ3           15.
    Z2Z11ZZ3  ACZ11ZZ3           15.   CCZ11ZZ3            1.
    Z3Z11ZZ3  OBJZZZZZ           25.   NCZZ6ZZ3           25.
    Z3Z11ZZ3  ACZ11ZZ3           25.   CCZ11ZZ3            1.
    Z3Z11ZZ3  NCZZ6ZZ4           -7.   ACZ11ZZ4           -7.
    Z4Z11ZZ3  OBJZZZZZ            1.   NCZZ6ZZ3            1.
    Z4Z11ZZ3  ACZ11ZZ3            1.   NCZZ6ZZ4           -1.
    Z4Z11ZZ3  ACZ11ZZ4           -1.
    Z1Z12ZZ3  CCZ12ZZ3            1.
    Z2Z12ZZ3  OBJZZZZZ           50.   NCZZ5ZZ3           50.
    Z2Z12ZZ3  ACZ12ZZ3           50.   CCZ12ZZ3            1.
    Z3Z12ZZ3  OBJZZZZZ           80.   NCZZ5ZZ3           80.
    Z3Z12ZZ3  ACZ12ZZ3           80.   CCZ12ZZ3            1.
    Z3Z12ZZ3  NCZZ5ZZ4          -15.   ACZ12ZZ4          -15.
    Z4Z12ZZ3  OBJZZZZZ            1.   NCZZ5ZZ3            1.
    Z4Z12ZZ3  ACZ12ZZ3            1.   NCZZ5ZZ4           -1.
    Z4Z12ZZ3  ACZ12ZZ4           -1.
    Z1ZZ1ZZ4  CCZZ1ZZ4            1.
    Z2ZZ1ZZ4  OBJZZZZZ           10.   NCZZ1ZZ4           10.
    Z2ZZ1ZZ4  ACZZ1ZZ4           10.   CCZZ1ZZ4            1.
    Z2ZZ1ZZ4  NCZZ2ZZ5          -10.
    Z3ZZ1ZZ4  OBJZZZZZ           20.   NCZZ1ZZ4           20.
    Z3ZZ1ZZ4  ACZZ1ZZ4           20.   CCZZ1ZZ4            1.
    Z3ZZ1ZZ4  NCZZ1ZZ5           -5.   ACZZ1ZZ5           -5.
    Z3ZZ1ZZ4  NCZZ2ZZ5          -15.
    Z4ZZ1ZZ4  OBJZZZZZ            1.   NCZZ1ZZ4            1.
    Z4ZZ1ZZ4  ACZZ1ZZ4            1.   NCZZ1ZZ5           -1.
    Z4ZZ1ZZ4  ACZZ1ZZ5           -1.
    Z1ZZ2ZZ4  CCZZ2ZZ4            1.
    Z2ZZ2ZZ4  OBJZZZZZ           15.   NCZZ2ZZ4           15.
    Z2ZZ2ZZ4  ACZZ2ZZ4           15.   CCZZ2ZZ4            1.
    Z2ZZ2ZZ4  NCZZ3ZZ5          -15.
    Z3ZZ2ZZ4  OBJZZZZZ           25.   NCZZ2ZZ4           25.
    Z3ZZ2ZZ4  ACZZ2ZZ4           25.   CCZZ2ZZ4            1.
    Z3ZZ2ZZ4  NCZZ2ZZ5           -7.   ACZZ2ZZ5           -7.
    Z3ZZ2ZZ4  NCZZ3ZZ5          -18.
    Z4ZZ2ZZ4  OBJZZZZZ            1.   NCZZ2ZZ4            1.
    Z4ZZ2ZZ4  ACZZ2ZZ4            1.   NCZZ2ZZ5           -1.
    Z4ZZ2ZZ4  ACZZ2ZZ5           -1.
    Z1ZZ3ZZ4  CCZZ3ZZ4            1.
    Z2ZZ3ZZ4  OBJZZZZZ           10.   NCZZ2ZZ4           10.
    Z2ZZ3ZZ4  ACZZ3ZZ4           10.   CCZZ3ZZ4            1.
    Z2ZZ3ZZ4  NCZZ1ZZ5          -10.
    Z3ZZ3ZZ4  OBJZZZZZ           20.   NCZZ2ZZ4           20.
    Z3ZZ3ZZ4  ACZZ3ZZ4           20.   CCZZ3ZZ4            1.
    Z3ZZ3ZZ4  NCZZ2ZZ5           -5.   ACZZ3ZZ5           -5.
    Z3ZZ3ZZ4  NCZZ1ZZ5          -15.
    Z4ZZ3ZZ4  OBJZZZZZ            1.   NCZZ2ZZ4            1.
    Z4ZZ3ZZ4  ACZZ3ZZ4            1.   NCZZ2ZZ5           -1.
    Z4ZZ3ZZ4  ACZZ3ZZ5           -1.
    Z1ZZ4ZZ4  CCZZ4ZZ4            1.
    Z2ZZ4ZZ4  OBJZZZZZ           15.   NCZZ1ZZ4           15.
    Z2ZZ4ZZ4  ACZZ4ZZ4           15.   CCZZ4ZZ4            1.
    Z2ZZ4ZZ4  NCZZ4ZZ5          -15.
    Z3ZZ4ZZ4  OBJZZZZZ           25.   NCZZ1ZZ4           25.
    Z3ZZ4ZZ4  ACZZ4ZZ4           25.   CCZZ4ZZ4            1.
    Z3ZZ4ZZ4  NCZZ1ZZ5           -7.   ACZZ4ZZ5           -7.
    Z3ZZ4ZZ4  NCZZ4ZZ5          -18.
    Z4ZZ4ZZ4  OBJZZZZZ            1.   NCZZ1ZZ4            1.
    Z4ZZ4ZZ4  ACZZ4ZZ4            1.   NCZZ1ZZ5           -1.
    Z4ZZ4ZZ4  ACZZ4ZZ5           -1.
    Z1ZZ5ZZ4  CCZZ5ZZ4            1.
    Z2ZZ5ZZ4  OBJZZZZZ           15.   NCZZ2ZZ4           15.
    Z2ZZ5ZZ4  ACZZ5ZZ4           15.   CCZZ5ZZ4            1.
    Z2ZZ5ZZ4  NCZZ4ZZ5          -15.
    Z3ZZ5ZZ4  OBJZZZZZ           25.   NCZZ2ZZ4           25.
    Z3ZZ5ZZ4  ACZZ5ZZ4           25.   CCZZ5ZZ4            1.
    Z3ZZ5ZZ4  NCZZ2ZZ5           -7.   ACZZ5ZZ5           -7.
    Z3ZZ5ZZ4  NCZZ4ZZ5          -18.
    Z4ZZ5ZZ4  OBJZZZZZ            1.   NCZZ2ZZ4            1.
    Z4ZZ5ZZ4  ACZZ5ZZ4            1.   NCZZ2ZZ5           -1.
    Z4ZZ5ZZ4  ACZZ5ZZ5           -1.
    Z1ZZ6ZZ4  CCZZ6ZZ4            1.
    Z2ZZ6ZZ4  OBJZZZZZ           10.   NCZZ2ZZ4           10.
    Z2ZZ6ZZ4  ACZZ6ZZ4           10.   CCZZ6ZZ4            1.
    Z2ZZ6ZZ4  NCZZ5ZZ5          -10.
    Z3ZZ6ZZ4  OBJZZZZZ           20.   NCZZ2ZZ4           20.
    Z3ZZ6ZZ4  ACZZ6ZZ4           20.   CCZZ6ZZ4            1.
    Z3ZZ6ZZ4  NCZZ2ZZ5           -5.   ACZZ6ZZ5           -5.
    Z3ZZ6ZZ4  NCZZ5ZZ5          -15.
    Z4ZZ6ZZ4  OBJZZZZZ            1.   NCZZ2ZZ4            1.
    Z4ZZ6ZZ4  ACZZ6ZZ4            1.   NCZZ2ZZ5           -1.
    Z4ZZ6ZZ4  ACZZ6ZZ5           -1.
    Z1ZZ7ZZ4  CCZZ7ZZ4            1.
    Z2ZZ7ZZ4  OBJZZZZZ           10.   NCZZ3ZZ4           10.
    Z2ZZ7ZZ4  ACZZ7ZZ4           10.   CCZZ7ZZ4            1.
    Z2ZZ7ZZ4  NCZZ5ZZ5          -10.
    Z3ZZ7ZZ4  OBJZZZZZ           20.   NCZZ3ZZ4           20.
    Z3ZZ7ZZ4  ACZZ7ZZ4           20.   CCZZ7ZZ4            1.
    Z3ZZ7ZZ4  NCZZ3ZZ5           -5.   ACZZ7ZZ5           -5.
    Z3ZZ7ZZ4  NCZZ5ZZ5          -15.
    Z4ZZ7ZZ4  OBJZZZZZ            1.   NCZZ3ZZ4            1.
    Z4ZZ7ZZ4  ACZZ7ZZ4            1.   NCZZ3ZZ5           -1.
    Z4ZZ7ZZ4  ACZZ7ZZ5           -1.
    Z1ZZ8ZZ4  CCZZ8ZZ4            1.
    Z2ZZ8ZZ4  OBJZZZZZ           30.   NCZZ4ZZ4           30.
    Z2ZZ8ZZ4  ACZZ8ZZ4           30.   CCZZ8ZZ4            1.
    Z2ZZ8ZZ4  NCZZ5ZZ5          -30.
    Z3ZZ8ZZ4  OBJZZZZZ           40.   NCZZ4ZZ4           40.
    Z3ZZ8ZZ4  ACZZ8ZZ4           40.   CCZZ8ZZ4            1.
    Z3ZZ8ZZ4  NCZZ4ZZ5           -7.   ACZZ8ZZ5           -7.
    Z3ZZ8ZZ4  NCZZ5ZZ5          -33.
    Z4ZZ8ZZ4  OBJZZZZZ            1.   NCZZ4ZZ4            1.
    Z4ZZ8ZZ4  ACZZ8ZZ4            1.   NCZZ4ZZ5           -1.
    Z4ZZ8ZZ4  ACZZ8ZZ5           -1.
    Z1ZZ9ZZ4  CCZZ9ZZ4            1.
    Z2ZZ9ZZ4  OBJZZZZZ           30.   NCZZ5ZZ4           30.
    Z2ZZ9ZZ4  ACZZ9ZZ4           30.   CCZZ9ZZ4            1.
    Z2ZZ9ZZ4  NCZZ4ZZ5          -30.
    Z3ZZ9ZZ4  OBJZZZZZ           40.   NCZZ5ZZ4           40.
    Z3ZZ9ZZ4  ACZZ9ZZ4           40.   CCZZ9ZZ4            1.
    Z3ZZ9ZZ4  NCZZ5ZZ5           -7.   ACZZ9ZZ5           -7.
    Z3ZZ9ZZ4  NCZZ4ZZ5          -33.
    Z4ZZ9ZZ4  OBJZZZZZ            1.   NCZZ5ZZ4            1.
    Z4ZZ9ZZ4  ACZZ9ZZ4            1.   NCZZ5ZZ5           -1.
    Z4ZZ9ZZ4  ACZZ9ZZ5           -1.
    Z1Z10ZZ4  CCZ10ZZ4            1.
    Z2Z10ZZ4  OBJZZZZZ           10.   NCZZ4ZZ4           10.
    Z2Z10ZZ4  ACZ10ZZ4           10.   CCZ10ZZ4            1.
    Z2Z10ZZ4  NCZZ6ZZ5          -10.
    Z3Z10ZZ4  OBJZZZZZ           20.   NCZZ4ZZ4           20.
    Z3Z10ZZ4  ACZ10ZZ4           20.   CCZ10ZZ4            1.
    Z3Z10ZZ4  NCZZ4ZZ5           -5.   ACZ10ZZ5           -5.
    Z3Z10ZZ4  NCZZ6ZZ5          -15.
    Z4Z10ZZ4  OBJZZZZZ            1.   NCZZ4ZZ4            1.
    Z4Z10ZZ4  ACZ10ZZ4            1.   NCZZ4ZZ5           -1.
    Z4Z10ZZ4  ACZ10ZZ5           -1.
    Z1Z11ZZ4  CCZ11ZZ4            1.
    Z2Z11ZZ4  OBJZZZZZ           15.   NCZZ6ZZ4           15.
    Z2Z11ZZ4  ACZ11ZZ4           15.   CCZ11ZZ4            1.
    Z3Z11ZZ4  OBJZZZZZ           25.   NCZZ6ZZ4           25.
    Z3Z11ZZ4  ACZ11ZZ4           25.   CCZ11ZZ4            1.
    Z3Z11ZZ4  NCZZ6ZZ5           -7.   ACZ11ZZ5           -7.
    Z4Z11ZZ4  OBJZZZZZ            1.   NCZZ6ZZ4            1.
    Z4Z11ZZ4  ACZ11ZZ4            1.   NCZZ6ZZ5           -1.
    Z4Z11ZZ4  ACZ11ZZ5           -1.
    Z1Z12ZZ4  CCZ12ZZ4            1.
    Z2Z12ZZ4  OBJZZZZZ           50.   NCZZ5ZZ4           50.
    Z2Z12ZZ4  ACZ12ZZ4           50.   CCZ12ZZ4            1.
    Z3Z12ZZ4  OBJZZZZZ           80.   NCZZ5ZZ4           80.
    Z3Z12ZZ4  ACZ12ZZ4           80.   CCZ12ZZ4            1.
    Z3Z12ZZ4  NCZZ5ZZ5          -15.   ACZ12ZZ5          -15.
    Z4Z12ZZ4  OBJZZZZZ            1.   NCZZ5ZZ4            1.
    Z4Z12ZZ4  ACZ12ZZ4            1.   NCZZ5ZZ5           -1.
    Z4Z12ZZ4  ACZ12ZZ5           -1.
    Z1ZZ1ZZ5  CCZZ1ZZ5            1.
    Z2ZZ1ZZ5  OBJZZZZZ           10.   NCZZ1ZZ5           10.
    Z2ZZ1ZZ5  ACZZ1ZZ5           10.   CCZZ1ZZ5            1.
    Z2ZZ1ZZ5  NCZZ2ZZ6          -10.
    Z3ZZ1ZZ5  OBJZZZZZ           20.   NCZZ1ZZ5           20.
    Z3ZZ1ZZ5  ACZZ1ZZ5           20.   CCZZ1ZZ5            1.
    Z3ZZ1ZZ5  NCZZ1ZZ6           -5.   ACZZ1ZZ6           -5.
    Z3ZZ1ZZ5  NCZZ2ZZ6          -15.
    Z4ZZ1ZZ5  OBJZZZZZ            1.   NCZZ1ZZ5            1.
    Z4ZZ1ZZ5  ACZZ1ZZ5            1.   NCZZ1ZZ6           -1.
    Z4ZZ1ZZ5  ACZZ1ZZ6           -1.
    Z1ZZ2ZZ5  CCZZ2ZZ5            1.
    Z2ZZ2ZZ5  OBJZZZZZ           15.   NCZZ2ZZ5           15.
    Z2ZZ2ZZ5  ACZZ2ZZ5           15.   CCZZ2ZZ5            1.
    Z2ZZ2ZZ5  NCZZ3ZZ6          -15.
    Z3ZZ2ZZ5  OBJZZZZZ           25.   NCZZ2ZZ5           25.
    Z3ZZ2ZZ5  ACZZ2ZZ5           25.   CCZZ2ZZ5            1.
    Z3ZZ2ZZ5  NCZZ2ZZ6           -7.   ACZZ2ZZ6           -7.
    Z3ZZ2ZZ5  NCZZ3ZZ6          -18.
    Z4ZZ2ZZ5  OBJZZZZZ            1.   NCZZ2ZZ5            1.
    Z4ZZ2ZZ5  ACZZ2ZZ5            1.   NCZZ2ZZ6           -1.
    Z4ZZ2ZZ5  ACZZ2ZZ6           -1.
    Z1ZZ3ZZ5  CCZZ3ZZ5            1.
    Z2ZZ3ZZ5  OBJZZZZZ           10.   NCZZ2ZZ5           10.
    Z2ZZ3ZZ5  ACZZ3ZZ5           10.   CCZZ3ZZ5            1.
    Z2ZZ3ZZ5  NCZZ1ZZ6          -10.
    Z3ZZ3ZZ5  OBJZZZZZ           20.   NCZZ2ZZ5           20.
    Z3ZZ3ZZ5  ACZZ3ZZ5           20.   CCZZ3ZZ5            1.
    Z3ZZ3ZZ5  NCZZ2ZZ6           -5.   ACZZ3ZZ6           -5.
    Z3ZZ3ZZ5  NCZZ1ZZ6          -15.
    Z4ZZ3ZZ5  OBJZZZZZ            1.   NCZZ2ZZ5            1.
    Z4ZZ3ZZ5  ACZZ3ZZ5            1.   NCZZ2ZZ6           -1.
    Z4ZZ3ZZ5  ACZZ3ZZ6           -1.
    Z1ZZ4ZZ5  CCZZ4ZZ5            1.
    Z2ZZ4ZZ5  OBJZZZZZ           15.   NCZZ1ZZ5           15.
    Z2ZZ4ZZ5  ACZZ4ZZ5           15.   CCZZ4ZZ5            1.
    Z2ZZ4ZZ5  NCZZ4ZZ6          -15.
    Z3ZZ4ZZ5  OBJZZZZZ           25.   NCZZ1ZZ5           25.
    Z3ZZ4ZZ5  ACZZ4ZZ5           25.   CCZZ4ZZ5            1.
    Z3ZZ4ZZ5  NCZZ1ZZ6           -7.   ACZZ4ZZ6           -7.
    Z3ZZ4ZZ5  NCZZ4ZZ6          -18.
    Z4ZZ4ZZ5  OBJZZZZZ            1.   NCZZ1ZZ5            1.
    Z4ZZ4ZZ5  ACZZ4ZZ5            1.   NCZZ1ZZ6           -1.
    Z4ZZ4ZZ5  ACZZ4ZZ6           -1.
    Z1ZZ5ZZ5  CCZZ5ZZ5            1.
    Z2ZZ5ZZ5  OBJZZZZZ           15.   NCZZ2ZZ5           15.
    Z2ZZ5ZZ5  ACZZ5ZZ5           15.   CCZZ5ZZ5            1.
    Z2ZZ5ZZ5  NCZZ4ZZ6          -15.
    Z3ZZ5ZZ5  OBJZZZZZ           25.   NCZZ2ZZ5           25.
    Z3ZZ5ZZ5  ACZZ5ZZ5           25.   CCZZ5ZZ5            1.
    Z3ZZ5ZZ5  NCZZ2ZZ6           -7.   ACZZ5ZZ6           -7.
    Z3ZZ5ZZ5  NCZZ4ZZ6          -18.
    Z4ZZ5ZZ5  OBJZZZZZ            1.   NCZZ2ZZ5            1.
    Z4ZZ5ZZ5  ACZZ5ZZ5            1.   NCZZ2ZZ6           -1.
    Z4ZZ5ZZ5  ACZZ5ZZ6           -1.
    Z1ZZ6ZZ5  CCZZ6ZZ5            1.
    Z2ZZ6ZZ5  OBJZZZZZ           10.   NCZZ2ZZ5           10.
    Z2ZZ6ZZ5  ACZZ6ZZ5           10.   CCZZ6ZZ5            1.
    Z2ZZ6ZZ5  NCZZ5ZZ6          -10.
    Z3ZZ6ZZ5  OBJZZZZZ           20.   NCZZ2ZZ5           20.
    Z3ZZ6ZZ5  ACZZ6ZZ5           20.   CCZZ6ZZ5            1.
    Z3ZZ6ZZ5  NCZZ2ZZ6           -5.   ACZZ6ZZ6           -5.
    Z3ZZ6ZZ5  NCZZ5ZZ6          -15.
    Z4ZZ6ZZ5  OBJZZZZZ            1.   NCZZ2ZZ5            1.
    Z4ZZ6ZZ5  ACZZ6ZZ5            1.   NCZZ2ZZ6           -1.
    Z4ZZ6ZZ5  ACZZ6ZZ6           -1.
    Z1ZZ7ZZ5  CCZZ7ZZ5            1.
    Z2ZZ7ZZ5  OBJZZZZZ           10.   NCZZ3ZZ5           10.
    Z2ZZ7ZZ5  ACZZ7ZZ5           10.   CCZZ7ZZ5            1.
    Z2ZZ7ZZ5  NCZZ5ZZ6          -10.
    Z3ZZ7ZZ5  OBJZZZZZ           20.   NCZZ3ZZ5           20.
    Z3ZZ7ZZ5  ACZZ7ZZ5           20.   CCZZ7ZZ5            1.
    Z3ZZ7ZZ5  NCZZ3ZZ6           -5.   ACZZ7ZZ6           -5.
    Z3ZZ7ZZ5  NCZZ5ZZ6          -15.
    Z4ZZ7ZZ5  OBJZZZZZ            1.   NCZZ3ZZ5            1.
    Z4ZZ7ZZ5  ACZZ7ZZ5            1.   NCZZ3ZZ6           -1.
    Z4ZZ7ZZ5  ACZZ7ZZ6           -1.
    Z1ZZ8ZZ5  CCZZ8ZZ5            1.
    Z2ZZ8ZZ5  OBJZZZZZ           30.   NCZZ4ZZ5           30.
    Z2ZZ8ZZ5  ACZZ8ZZ5           30.   CCZZ8ZZ5            1.
    Z2ZZ8ZZ5  NCZZ5ZZ6          -30.
    Z3ZZ8ZZ5  OBJZZZZZ           40.   NCZZ4ZZ5           40.
    Z3ZZ8ZZ5  ACZZ8ZZ5           40.   CCZZ8ZZ5            1.
    Z3ZZ8ZZ5  NCZZ4ZZ6           -7.   ACZZ8ZZ6           -7.
    Z3ZZ8ZZ5  NCZZ5ZZ6          -33.
    Z4ZZ8ZZ5  OBJZZZZZ            1.   NCZZ4ZZ5            1.
    Z4ZZ8ZZ5  ACZZ8ZZ5            1.   NCZZ4ZZ6           -1.
    Z4ZZ8ZZ5  ACZZ8ZZ6           -1.
    Z1ZZ9ZZ5  CCZZ9ZZ5            1.
    Z2ZZ9ZZ5  OBJZZZZZ           30.   NCZZ5ZZ5           30.
    Z2ZZ9ZZ5  ACZZ9ZZ5           30.   CCZZ9ZZ5            1.
    Z2ZZ9ZZ5  NCZZ4ZZ6          -30.
    Z3ZZ9ZZ5  OBJZZZZZ           40.   NCZZ5ZZ5           40.
    Z3ZZ9ZZ5  ACZZ9ZZ5           40.   CCZZ9ZZ5            1.
    Z3ZZ9ZZ5  NCZZ5ZZ6           -7.   ACZZ9ZZ6           -7.
    Z3ZZ9ZZ5  NCZZ4ZZ6          -33.
    Z4ZZ9ZZ5  OBJZZZZZ            1.   NCZZ5ZZ5            1.
    Z4ZZ9ZZ5  ACZZ9ZZ5            1.   NCZZ5ZZ6           -1.
    Z4ZZ9ZZ5  ACZZ9ZZ6           -1.
    Z1Z10ZZ5  CCZ10ZZ5            1.
    Z2Z10ZZ5  OBJZZZZZ           10.   NCZZ4ZZ5           10.
    Z2Z10ZZ5  ACZ10ZZ5           10.   CCZ10ZZ5            1.
    Z2Z10ZZ5  NCZZ6ZZ6          -10.
    Z3Z10ZZ5  OBJZZZZZ           20.   NCZZ4ZZ5           20.
    Z3Z10ZZ5  ACZ10ZZ5           20.   CCZ10ZZ5            1.
    Z3Z10ZZ5  NCZZ4ZZ6           -5.   ACZ10ZZ6           -5.
    Z3Z10ZZ5  NCZZ6ZZ6          -15.
    Z4Z10ZZ5  OBJZZZZZ            1.   NCZZ4ZZ5            1.
    Z4Z10ZZ5  ACZ10ZZ5            1.   NCZZ4ZZ6           -1.
    Z4Z10ZZ5  ACZ10ZZ6           -1.
    Z1Z11ZZ5  CCZ11ZZ5            1.
    Z2Z11ZZ5  OBJZZZZZ           15.   NCZZ6ZZ5           15.
    Z2Z11ZZ5  ACZ11ZZ5           15.   CCZ11ZZ5            1.
    Z3Z11ZZ5  OBJZZZZZ           25.   NCZZ6ZZ5           25.
    Z3Z11ZZ5  ACZ11ZZ5           25.   CCZ11ZZ5            1.
    Z3Z11ZZ5  NCZZ6ZZ6           -7.   ACZ11ZZ6           -7.
    Z4Z11ZZ5  OBJZZZZZ            1.   NCZZ6ZZ5            1.
    Z4Z11ZZ5  ACZ11ZZ5            1.   NCZZ6ZZ6           -1.
    Z4Z11ZZ5  ACZ11ZZ6           -1.
    Z1Z12ZZ5  CCZ12ZZ5            1.
    Z2Z12ZZ5  OBJZZZZZ           50.   NCZZ5ZZ5           50.
    Z2Z12ZZ5  ACZ12ZZ5           50.   CCZ12ZZ5            1.
    Z3Z12ZZ5  OBJZZZZZ           80.   NCZZ5ZZ5           80.
    Z3Z12ZZ5  ACZ12ZZ5           80.   CCZ12ZZ5            1.
    Z3Z12ZZ5  NCZZ5ZZ6          -15.   ACZ12ZZ6          -15.
    Z4Z12ZZ5  OBJZZZZZ            1.   NCZZ5ZZ5            1.
    Z4Z12ZZ5  ACZ12ZZ5            1.   NCZZ5ZZ6           -1.
    Z4Z12ZZ5  ACZ12ZZ6           -1.
    Z1ZZ1ZZ6  CCZZ1ZZ6            1.
    Z2ZZ1ZZ6  OBJZZZZZ           10.   NCZZ1ZZ6           10.
    Z2ZZ1ZZ6  ACZZ1ZZ6           10.   CCZZ1ZZ6            1.
    Z2ZZ1ZZ6  NCZZ2ZZ7          -10.
    Z3ZZ1ZZ6  OBJZZZZZ           20.   NCZZ1ZZ6           20.
    Z3ZZ1ZZ6  ACZZ1ZZ6           20.   CCZZ1ZZ6            1.
    Z3ZZ1ZZ6  NCZZ1ZZ7           -5.   ACZZ1ZZ7           -5.
    Z3ZZ1ZZ6  NCZZ2ZZ7          -15.
    Z4ZZ1ZZ6  OBJZZZZZ            1.   NCZZ1ZZ6            1.
    Z4ZZ1ZZ6  ACZZ1ZZ6            1.   NCZZ1ZZ7           -1.
    Z4ZZ1ZZ6  ACZZ1ZZ7           -1.
    Z1ZZ2ZZ6  CCZZ2ZZ6            1.
    Z2ZZ2ZZ6  OBJZZZZZ           15.   NCZZ2ZZ6           15.
    Z2ZZ2ZZ6  ACZZ2ZZ6           15.   CCZZ2ZZ6            1.
    Z2ZZ2ZZ6  NCZZ3ZZ7          -15.
    Z3ZZ2ZZ6  OBJZZZZZ           25.   NCZZ2ZZ6           25.
    Z3ZZ2ZZ6  ACZZ2ZZ6           25.   CCZZ2ZZ6            1.
    Z3ZZ2ZZ6  NCZZ2ZZ7           -7.   ACZZ2ZZ7           -7.
    Z3ZZ2ZZ6  NCZZ3ZZ7          -18.
    Z4ZZ2ZZ6  OBJZZZZZ            1.   NCZZ2ZZ6            1.
    Z4ZZ2ZZ6  ACZZ2ZZ6            1.   NCZZ2ZZ7           -1.
    Z4ZZ2ZZ6  ACZZ2ZZ7           -1.
    Z1ZZ3ZZ6  CCZZ3ZZ6            1.
    Z2ZZ3ZZ6  OBJZZZZZ           10.   NCZZ2ZZ6           10.
    Z2ZZ3ZZ6  ACZZ3ZZ6           10.   CCZZ3ZZ6            1.
    Z2ZZ3ZZ6  NCZZ1ZZ7          -10.
    Z3ZZ3ZZ6  OBJZZZZZ           20.   NCZZ2ZZ6           20.
    Z3ZZ3ZZ6  ACZZ3ZZ6           20.   CCZZ3ZZ6            1.
    Z3ZZ3ZZ6  NCZZ2ZZ7           -5.   ACZZ3ZZ7           -5.
    Z3ZZ3ZZ6  NCZZ1ZZ7          -15.
    Z4ZZ3ZZ6  OBJZZZZZ            1.   NCZZ2ZZ6            1.
    Z4ZZ3ZZ6  ACZZ3ZZ6            1.   NCZZ2ZZ7           -1.
    Z4ZZ3ZZ6  ACZZ3ZZ7           -1.
    Z1ZZ4ZZ6  CCZZ4ZZ6            1.
    Z2ZZ4ZZ6  OBJZZZZZ           15.   NCZZ1ZZ6           15.
    Z2ZZ4ZZ6  ACZZ4ZZ6           15.   CCZZ4ZZ6            1.
    Z2ZZ4ZZ6  NCZZ4ZZ7          -15.
    Z3ZZ4ZZ6  OBJZZZZZ           25.   NCZZ1ZZ6           25.
    Z3ZZ4ZZ6  ACZZ4ZZ6           25.   CCZZ4ZZ6            1.
    Z3ZZ4ZZ6  NCZZ1ZZ7           -7.   ACZZ4ZZ7           -7.
    Z3ZZ4ZZ6  NCZZ4ZZ7          -18.
    Z4ZZ4ZZ6  OBJZZZZZ            1.   NCZZ1ZZ6            1.
    Z4ZZ4ZZ6  ACZZ4ZZ6            1.   NCZZ1ZZ7           -1.
    Z4ZZ4ZZ6  ACZZ4ZZ7           -1.
    Z1ZZ5ZZ6  CCZZ5ZZ6            1.
    Z2ZZ5ZZ6  OBJZZZZZ           15.   NCZZ2ZZ6           15.
    Z2ZZ5ZZ6  ACZZ5ZZ6           15.   CCZZ5ZZ6            1.
    Z2ZZ5ZZ6  NCZZ4ZZ7          -15.
    Z3ZZ5ZZ6  OBJZZZZZ           25.   NCZZ2ZZ6           25.
    Z3ZZ5ZZ6  ACZZ5ZZ6           25.   CCZZ5ZZ6            1.
    Z3ZZ5ZZ6  NCZZ2ZZ7           -7.   ACZZ5ZZ7           -7.
    Z3ZZ5ZZ6  NCZZ4ZZ7          -18.
    Z4ZZ5ZZ6  OBJZZZZZ            1.   NCZZ2ZZ6            1.
    Z4ZZ5ZZ6  ACZZ5ZZ6            1.   NCZZ2ZZ7           -1.
    Z4ZZ5ZZ6  ACZZ5ZZ7           -1.
    Z1ZZ6ZZ6  CCZZ6ZZ6            1.
    Z2ZZ6ZZ6  OBJZZZZZ           10.   NCZZ2ZZ6           10.
    Z2ZZ6ZZ6  ACZZ6ZZ6           10.   CCZZ6ZZ6            1.
    Z2ZZ6ZZ6  NCZZ5ZZ7          -10.
    Z3ZZ6ZZ6  OBJZZZZZ           20.   NCZZ2ZZ6           20.
    Z3ZZ6ZZ6  ACZZ6ZZ6           20.   CCZZ6ZZ6            1.
    Z3ZZ6ZZ6  NCZZ2ZZ7           -5.   ACZZ6ZZ7           -5.
    Z3ZZ6ZZ6  NCZZ5ZZ7          -15.
    Z4ZZ6ZZ6  OBJZZZZZ            1.   NCZZ2ZZ6            1.
    Z4ZZ6ZZ6  ACZZ6ZZ6            1.   NCZZ2ZZ7           -1.
    Z4ZZ6ZZ6  ACZZ6ZZ7           -1.
    Z1ZZ7ZZ6  CCZZ7ZZ6            1.
    Z2ZZ7ZZ6  OBJZZZZZ           10.   NCZZ3ZZ6           10.
    Z2ZZ7ZZ6  ACZZ7ZZ6           10.   CCZZ7ZZ6            1.
    Z2ZZ7ZZ6  NCZZ5ZZ7          -10.
    Z3ZZ7ZZ6  OBJZZZZZ           20.   NCZZ3ZZ6           20.
    Z3ZZ7ZZ6  ACZZ7ZZ6           20.   CCZZ7ZZ6            1.
    Z3ZZ7ZZ6  NCZZ3ZZ7           -5.   ACZZ7ZZ7           -5.
    Z3ZZ7ZZ6  NCZZ5ZZ7          -15.
    Z4ZZ7ZZ6  OBJZZZZZ            1.   NCZZ3ZZ6            1.
    Z4ZZ7ZZ6  ACZZ7ZZ6            1.   NCZZ3ZZ7           -1.
    Z4ZZ7ZZ6  ACZZ7ZZ7           -1.
    Z1ZZ8ZZ6  CCZZ8ZZ6            1.
    Z2ZZ8ZZ6  OBJZZZZZ           30.   NCZZ4ZZ6           30.
    Z2ZZ8ZZ6  ACZZ8ZZ6           30.   CCZZ8ZZ6            1.
    Z2ZZ8ZZ6  NCZZ5ZZ7          -30.
    Z3ZZ8ZZ6  OBJZZZZZ           40.   NCZZ4ZZ6           40.
    Z3ZZ8ZZ6  ACZZ8ZZ6           40.   CCZZ8ZZ6            1.
    Z3ZZ8ZZ6  NCZZ4ZZ7           -7.   ACZZ8ZZ7           -7.
    Z3ZZ8ZZ6  NCZZ5ZZ7          -33.
    Z4ZZ8ZZ6  OBJZZZZZ            1.   NCZZ4ZZ6            1.
    Z4ZZ8ZZ6  ACZZ8ZZ6            1.   NCZZ4ZZ7           -1.
    Z4ZZ8ZZ6  ACZZ8ZZ7           -1.
    Z1ZZ9ZZ6  CCZZ9ZZ6            1.
    Z2ZZ9ZZ6  OBJZZZZZ           30.   NCZZ5ZZ6           30.
    Z2ZZ9ZZ6  ACZZ9ZZ6           30.   CCZZ9ZZ6            1.
    Z2ZZ9ZZ6  NCZZ4ZZ7          -30.
    Z3ZZ9ZZ6  OBJZZZZZ           40.   NCZZ5ZZ6           40.
    Z3ZZ9ZZ6  ACZZ9ZZ6           40.   CCZZ9ZZ6            1.
    Z3ZZ9ZZ6  NCZZ5ZZ7           -7.   ACZZ9ZZ7           -7.
    Z3ZZ9ZZ6  NCZZ4ZZ7          -33.
    Z4ZZ9ZZ6  OBJZZZZZ            1.   NCZZ5ZZ6            1.
    Z4ZZ9ZZ6  ACZZ9ZZ6            1.   NCZZ5ZZ7           -1.
    Z4ZZ9ZZ6  ACZZ9ZZ7           -1.
    Z1Z10ZZ6  CCZ10ZZ6            1.
    Z2Z10ZZ6  OBJZZZZZ           10.   NCZZ4ZZ6           10.
    Z2Z10ZZ6  ACZ10ZZ6           10.   CCZ10ZZ6            1.
    Z2Z10ZZ6  NCZZ6ZZ7          -10.
    Z3Z10ZZ6  OBJZZZZZ           20.   NCZZ4ZZ6           20.
    Z3Z10ZZ6  ACZ10ZZ6           20.   CCZ10ZZ6            1.
    Z3Z10ZZ6  NCZZ4ZZ7           -5.   ACZ10ZZ7           -5.
    Z3Z10ZZ6  NCZZ6ZZ7          -15.
    Z4Z10ZZ6  OBJZZZZZ            1.   NCZZ4ZZ6            1.
    Z4Z10ZZ6  ACZ10ZZ6            1.   NCZZ4ZZ7           -1.
    Z4Z10ZZ6  ACZ10ZZ7           -1.
    Z1Z11ZZ6  CCZ11ZZ6            1.
    Z2Z11ZZ6  OBJZZZZZ           15.   NCZZ6ZZ6           15.
    Z2Z11ZZ6  ACZ11ZZ6           15.   CCZ11ZZ6            1.
    Z3Z11ZZ6  OBJZZZZZ           25.   NCZZ6ZZ6           25.
    Z3Z11ZZ6  ACZ11ZZ6           25.   CCZ11ZZ6            1.
    Z3Z11ZZ6  NCZZ6ZZ7           -7.   ACZ11ZZ7           -7.
    Z4Z11ZZ6  OBJZZZZZ            1.   NCZZ6ZZ6            1.
    Z4Z11ZZ6  ACZ11ZZ6            1.   NCZZ6ZZ7           -1.
    Z4Z11ZZ6  ACZ11ZZ7           -1.
    Z1Z12ZZ6  CCZ12ZZ6            1.
    Z2Z12ZZ6  OBJZZZZZ           50.   NCZZ5ZZ6           50.
    Z2Z12ZZ6  ACZ12ZZ6           50.   CCZ12ZZ6            1.
    Z3Z12ZZ6  OBJZZZZZ           80.   NCZZ5ZZ6           80.
    Z3Z12ZZ6  ACZ12ZZ6           80.   CCZ12ZZ6            1.
    Z3Z12ZZ6  NCZZ5ZZ7          -15.   ACZ12ZZ7          -15.
    Z4Z12ZZ6  OBJZZZZZ            1.   NCZZ5ZZ6            1.
    Z4Z12ZZ6  ACZ12ZZ6            1.   NCZZ5ZZ7           -1.
    Z4Z12ZZ6  ACZ12ZZ7           -1.
    Z1ZZ1ZZ7  CCZZ1ZZ7            1.
    Z2ZZ1ZZ7  OBJZZZZZ           10.   NCZZ1ZZ7           10.
    Z2ZZ1ZZ7  ACZZ1ZZ7           10.   CCZZ1ZZ7            1.
    Z2ZZ1ZZ7  NCZZ2ZZ8          -10.
    Z3ZZ1ZZ7  OBJZZZZZ           20.   NCZZ1ZZ7           20.
    Z3ZZ1ZZ7  ACZZ1ZZ7           20.   CCZZ1ZZ7            1.
    Z3ZZ1ZZ7  NCZZ1ZZ8           -5.   ACZZ1ZZ8           -5.
    Z3ZZ1ZZ7  NCZZ2ZZ8          -15.
    Z4ZZ1ZZ7  OBJZZZZZ            1.   NCZZ1ZZ7            1.
    Z4ZZ1ZZ7  ACZZ1ZZ7            1.   NCZZ1ZZ8           -1.
    Z4ZZ1ZZ7  ACZZ1ZZ8           -1.
    Z1ZZ2ZZ7  CCZZ2ZZ7            1.
    Z2ZZ2ZZ7  OBJZZZZZ           15.   NCZZ2ZZ7           15.
    Z2ZZ2ZZ7  ACZZ2ZZ7           15.   CCZZ2ZZ7            1.
    Z2ZZ2ZZ7  NCZZ3ZZ8          -15.
    Z3ZZ2ZZ7  OBJZZZZZ           25.   NCZZ2ZZ7           25.
    Z3ZZ2ZZ7  ACZZ2ZZ7           25.   CCZZ2ZZ7            1.
    Z3ZZ2ZZ7  NCZZ2ZZ8           -7.   ACZZ2ZZ8           -7.
    Z3ZZ2ZZ7  NCZZ3ZZ8          -18.
    Z4ZZ2ZZ7  OBJZZZZZ            1.   NCZZ2ZZ7            1.
    Z4ZZ2ZZ7  ACZZ2ZZ7            1.   NCZZ2ZZ8           -1.
    Z4ZZ2ZZ7  ACZZ2ZZ8           -1.
    Z1ZZ3ZZ7  CCZZ3ZZ7            1.
    Z2ZZ3ZZ7  OBJZZZZZ           10.   NCZZ2ZZ7           10.
    Z2ZZ3ZZ7  ACZZ3ZZ7           10.   CCZZ3ZZ7            1.
    Z2ZZ3ZZ7  NCZZ1ZZ8          -10.
    Z3ZZ3ZZ7  OBJZZZZZ           20.   NCZZ2ZZ7           20.
    Z3ZZ3ZZ7  ACZZ3ZZ7           20.   CCZZ3ZZ7            1.
    Z3ZZ3ZZ7  NCZZ2ZZ8           -5.   ACZZ3ZZ8           -5.
    Z3ZZ3ZZ7  NCZZ1ZZ8          -15.
    Z4ZZ3ZZ7  OBJZZZZZ            1.   NCZZ2ZZ7            1.
    Z4ZZ3ZZ7  ACZZ3ZZ7            1.   NCZZ2ZZ8           -1.
    Z4ZZ3ZZ7  ACZZ3ZZ8           -1.
    Z1ZZ4ZZ7  CCZZ4ZZ7            1.
    Z2ZZ4ZZ7  OBJZZZZZ           15.   NCZZ1ZZ7           15.
    Z2ZZ4ZZ7  ACZZ4ZZ7           15.   CCZZ4ZZ7            1.
    Z2ZZ4ZZ7  NCZZ4ZZ8          -15.
    Z3ZZ4ZZ7  OBJZZZZZ           25.   NCZZ1ZZ7           25.
    Z3ZZ4ZZ7  ACZZ4ZZ7           25.   CCZZ4ZZ7            1.
    Z3ZZ4ZZ7  NCZZ1ZZ8           -7.   ACZZ4ZZ8           -7.
    Z3ZZ4ZZ7  NCZZ4ZZ8          -18.
    Z4ZZ4ZZ7  OBJZZZZZ            1.   NCZZ1ZZ7            1.
    Z4ZZ4ZZ7  ACZZ4ZZ7            1.   NCZZ1ZZ8           -1.
    Z4ZZ4ZZ7  ACZZ4ZZ8           -1.
    Z1ZZ5ZZ7  CCZZ5ZZ7            1.
    Z2ZZ5ZZ7  OBJZZZZZ           15.   NCZZ2ZZ7           15.
    Z2ZZ5ZZ7  ACZZ5ZZ7           15.   CCZZ5ZZ7            1.
    Z2ZZ5ZZ7  NCZZ4ZZ8          -15.
    Z3ZZ5ZZ7  OBJZZZZZ           25.   NCZZ2ZZ7           25.
    Z3ZZ5ZZ7  ACZZ5ZZ7           25.   CCZZ5ZZ7            1.
    Z3ZZ5ZZ7  NCZZ2ZZ8           -7.   ACZZ5ZZ8           -7.
    Z3ZZ5ZZ7  NCZZ4ZZ8          -18.
    Z4ZZ5ZZ7  OBJZZZZZ            1.   NCZZ2ZZ7            1.
    Z4ZZ5ZZ7  ACZZ5ZZ7            1.   NCZZ2ZZ8           -1.
    Z4ZZ5ZZ7  ACZZ5ZZ8           -1.
    Z1ZZ6ZZ7  CCZZ6ZZ7            1.
    Z2ZZ6ZZ7  OBJZZZZZ           10.   NCZZ2ZZ7           10.
    Z2ZZ6ZZ7  ACZZ6ZZ7           10.   CCZZ6ZZ7            1.
    Z2ZZ6ZZ7  NCZZ5ZZ8          -10.
    Z3ZZ6ZZ7  OBJZZZZZ           20.   NCZZ2ZZ7           20.
    Z3ZZ6ZZ7  ACZZ6ZZ7           20.   CCZZ6ZZ7            1.
    Z3ZZ6ZZ7  NCZZ2ZZ8           -5.   ACZZ6ZZ8           -5.
    Z3ZZ6ZZ7  NCZZ5ZZ8          -15.
    Z4ZZ6ZZ7  OBJZZZZZ            1.   NCZZ2ZZ7            1.
    Z4ZZ6ZZ7  ACZZ6ZZ7            1.   NCZZ2ZZ8           -1.
    Z4ZZ6ZZ7  ACZZ6ZZ8           -1.
    Z1ZZ7ZZ7  CCZZ7ZZ7            1.
    Z2ZZ7ZZ7  OBJZZZZZ           10.   NCZZ3ZZ7           10.
    Z2ZZ7ZZ7  ACZZ7ZZ7           10.   CCZZ7ZZ7            1.
    Z2ZZ7ZZ7  NCZZ5ZZ8          -10.
    Z3ZZ7ZZ7  OBJZZZZZ           20.   NCZZ3ZZ7           20.
    Z3ZZ7ZZ7  ACZZ7ZZ7           20.   CCZZ7ZZ7            1.
    Z3ZZ7ZZ7  NCZZ3ZZ8           -5.   ACZZ7ZZ8           -5.
    Z3ZZ7ZZ7  NCZZ5ZZ8          -15.
    Z4ZZ7ZZ7  OBJZZZZZ            1.   NCZZ3ZZ7            1.
    Z4ZZ7ZZ7  ACZZ7ZZ7            1.   NCZZ3ZZ8           -1.
    Z4ZZ7ZZ7  ACZZ7ZZ8           -1.
    Z1ZZ8ZZ7  CCZZ8ZZ7            1.
    Z2ZZ8ZZ7  OBJZZZZZ           30.   NCZZ4ZZ7           30.
    Z2ZZ8ZZ7  ACZZ8ZZ7           30.   CCZZ8ZZ7            1.
    Z2ZZ8ZZ7  NCZZ5ZZ8          -30.
    Z3ZZ8ZZ7  OBJZZZZZ           40.   NCZZ4ZZ7           40.
    Z3ZZ8ZZ7  ACZZ8ZZ7           40.   CCZZ8ZZ7            1.
    Z3ZZ8ZZ7  NCZZ4ZZ8           -7.   ACZZ8ZZ8           -7.
    Z3ZZ8ZZ7  NCZZ5ZZ8          -33.
    Z4ZZ8ZZ7  OBJZZZZZ            1.   NCZZ4ZZ7            1.
    Z4ZZ8ZZ7  ACZZ8ZZ7            1.   NCZZ4ZZ8           -1.
    Z4ZZ8ZZ7  ACZZ8ZZ8           -1.
    Z1ZZ9ZZ7  CCZZ9ZZ7            1.
    Z2ZZ9ZZ7  OBJZZZZZ           30.   NCZZ5ZZ7           30.
    Z2ZZ9ZZ7  ACZZ9ZZ7           30.   CCZZ9ZZ7            1.
    Z2ZZ9ZZ7  NCZZ4ZZ8          -30.
    Z3ZZ9ZZ7  OBJZZZZZ           40.   NCZZ5ZZ7           40.
    Z3ZZ9ZZ7  ACZZ9ZZ7           40.   CCZZ9ZZ7            1.
    Z3ZZ9ZZ7  NCZZ5ZZ8           -7.   ACZZ9ZZ8           -7.
    Z3ZZ9ZZ7  NCZZ4ZZ8          -33.
    Z4ZZ9ZZ7  OBJZZZZZ            1.   NCZZ5ZZ7            1.
    Z4ZZ9ZZ7  ACZZ9ZZ7            1.   NCZZ5ZZ8           -1.
    Z4ZZ9ZZ7  ACZZ9ZZ8           -1.
    Z1Z10ZZ7  CCZ10ZZ7            1.
    Z2Z10ZZ7  OBJZZZZZ           10.   NCZZ4ZZ7           10.
    Z2Z10ZZ7  ACZ10ZZ7           10.   CCZ10ZZ7            1.
    Z2Z10ZZ7  NCZZ6ZZ8          -10.
    Z3Z10ZZ7  OBJZZZZZ           20.   NCZZ4ZZ7           20.
    Z3Z10ZZ7  ACZ10ZZ7           20.   CCZ10ZZ7            1.
    Z3Z10ZZ7  NCZZ4ZZ8           -5.   ACZ10ZZ8           -5.
    Z3Z10ZZ7  NCZZ6ZZ8          -15.
    Z4Z10ZZ7  OBJZZZZZ            1.   NCZZ4ZZ7            1.
    Z4Z10ZZ7  ACZ10ZZ7            1.   NCZZ4ZZ8           -1.
    Z4Z10ZZ7  ACZ10ZZ8           -1.
    Z1Z11ZZ7  CCZ11ZZ7            1.
    Z2Z11ZZ7  OBJZZZZZ           15.   NCZZ6ZZ7           15.
    Z2Z11ZZ7  ACZ11ZZ7           15.   CCZ11ZZ7            1.
    Z3Z11ZZ7  OBJZZZZZ           25.   NCZZ6ZZ7           25.
    Z3Z11ZZ7  ACZ11ZZ7           25.   CCZ11ZZ7            1.
    Z3Z11ZZ7  NCZZ6ZZ8           -7.   ACZ11ZZ8           -7.
    Z4Z11ZZ7  OBJZZZZZ            1.   NCZZ6ZZ7            1.
    Z4Z11ZZ7  ACZ11ZZ7            1.   NCZZ6ZZ8           -1.
    Z4Z11ZZ7  ACZ11ZZ8           -1.
    Z1Z12ZZ7  CCZ12ZZ7            1.
    Z2Z12ZZ7  OBJZZZZZ           50.   NCZZ5ZZ7           50.
    Z2Z12ZZ7  ACZ12ZZ7           50.   CCZ12ZZ7            1.
    Z3Z12ZZ7  OBJZZZZZ           80.   NCZZ5ZZ7           80.
    Z3Z12ZZ7  ACZ12ZZ7           80.   CCZ12ZZ7            1.
    Z3Z12ZZ7  NCZZ5ZZ8          -15.   ACZ12ZZ8          -15.
    Z4Z12ZZ7  OBJZZZZZ            1.   NCZZ5ZZ7            1.
    Z4Z12ZZ7  ACZ12ZZ7            1.   NCZZ5ZZ8           -1.
    Z4Z12ZZ7  ACZ12ZZ8           -1.
    Z1ZZ1ZZ8  CCZZ1ZZ8            1.
    Z2ZZ1ZZ8  OBJZZZZZ           10.   NCZZ1ZZ8           10.
    Z2ZZ1ZZ8  ACZZ1ZZ8           10.   CCZZ1ZZ8            1.
    Z2ZZ1ZZ8  NCZZ2ZZ9          -10.
    Z3ZZ1ZZ8  OBJZZZZZ           20.   NCZZ1ZZ8           20.
    Z3ZZ1ZZ8  ACZZ1ZZ8           20.   CCZZ1ZZ8            1.
    Z3ZZ1ZZ8  NCZZ1ZZ9           -5.   ACZZ1ZZ9           -5.
    Z3ZZ1ZZ8  NCZZ2ZZ9          -15.
    Z4ZZ1ZZ8  OBJZZZZZ            1.   NCZZ1ZZ8            1.
    Z4ZZ1ZZ8  ACZZ1ZZ8            1.   NCZZ1ZZ9           -1.
    Z4ZZ1ZZ8  ACZZ1ZZ9           -1.
    Z1ZZ2ZZ8  CCZZ2ZZ8            1.
    Z2ZZ2ZZ8  OBJZZZZZ           15.   NCZZ2ZZ8           15.
    Z2ZZ2ZZ8  ACZZ2ZZ8           15.   CCZZ2ZZ8            1.
    Z2ZZ2ZZ8  NCZZ3ZZ9          -15.
    Z3ZZ2ZZ8  OBJZZZZZ           25.   NCZZ2ZZ8           25.
    Z3ZZ2ZZ8  ACZZ2ZZ8           25.   CCZZ2ZZ8            1.
    Z3ZZ2ZZ8  NCZZ2ZZ9           -7.   ACZZ2ZZ9           -7.
    Z3ZZ2ZZ8  NCZZ3ZZ9          -18.
    Z4ZZ2ZZ8  OBJZZZZZ            1.   NCZZ2ZZ8            1.
    Z4ZZ2ZZ8  ACZZ2ZZ8            1.   NCZZ2ZZ9           -1.
    Z4ZZ2ZZ8  ACZZ2ZZ9           -1.
    Z1ZZ3ZZ8  CCZZ3ZZ8            1.
    Z2ZZ3ZZ8  OBJZZZZZ           10.   NCZZ2ZZ8           10.
    Z2ZZ3ZZ8  ACZZ3ZZ8           10.   CCZZ3ZZ8            1.
    Z2ZZ3ZZ8  NCZZ1ZZ9          -10.
    Z3ZZ3ZZ8  OBJZZZZZ           20.   NCZZ2ZZ8           20.
    Z3ZZ3ZZ8  ACZZ3ZZ8           20.   CCZZ3ZZ8            1.
    Z3ZZ3ZZ8  NCZZ2ZZ9           -5.   ACZZ3ZZ9           -5.
    Z3ZZ3ZZ8  NCZZ1ZZ9          -15.
    Z4ZZ3ZZ8  OBJZZZZZ            1.   NCZZ2ZZ8            1.
    Z4ZZ3ZZ8  ACZZ3ZZ8            1.   NCZZ2ZZ9           -1.
    Z4ZZ3ZZ8  ACZZ3ZZ9           -1.
    Z1ZZ4ZZ8  CCZZ4ZZ8            1.
    Z2ZZ4ZZ8  OBJZZZZZ           15.   NCZZ1ZZ8           15.
    Z2ZZ4ZZ8  ACZZ4ZZ8           15.   CCZZ4ZZ8            1.
    Z2ZZ4ZZ8  NCZZ4ZZ9          -15.
    Z3ZZ4ZZ8  OBJZZZZZ           25.   NCZZ1ZZ8           25.
    Z3ZZ4ZZ8  ACZZ4ZZ8           25.   CCZZ4ZZ8            1.
    Z3ZZ4ZZ8  NCZZ1ZZ9           -7.   ACZZ4ZZ9           -7.
    Z3ZZ4ZZ8  NCZZ4ZZ9          -18.
    Z4ZZ4ZZ8  OBJZZZZZ            1.   NCZZ1ZZ8            1.
    Z4ZZ4ZZ8  ACZZ4ZZ8            1.   NCZZ1ZZ9           -1.
    Z4ZZ4ZZ8  ACZZ4ZZ9           -1.
    Z1ZZ5ZZ8  CCZZ5ZZ8            1.
    Z2ZZ5ZZ8  OBJZZZZZ           15.   NCZZ2ZZ8           15.
    Z2ZZ5ZZ8  ACZZ5ZZ8           15.   CCZZ5ZZ8            1.
    Z2ZZ5ZZ8  NCZZ4ZZ9          -15.
    Z3ZZ5ZZ8  OBJZZZZZ           25.   NCZZ2ZZ8           25.
    Z3ZZ5ZZ8  ACZZ5ZZ8           25.   CCZZ5ZZ8            1.
    Z3ZZ5ZZ8  NCZZ2ZZ9           -7.   ACZZ5ZZ9           -7.
    Z3ZZ5ZZ8  NCZZ4ZZ9          -18.
    Z4ZZ5ZZ8  OBJZZZZZ            1.   NCZZ2ZZ8            1.
    Z4ZZ5ZZ8  ACZZ5ZZ8            1.   NCZZ2ZZ9           -1.
    Z4ZZ5ZZ8  ACZZ5ZZ9           -1.
    Z1ZZ6ZZ8  CCZZ6ZZ8            1.
    Z2ZZ6ZZ8  OBJZZZZZ           10.   NCZZ2ZZ8           10.
    Z2ZZ6ZZ8  ACZZ6ZZ8           10.   CCZZ6ZZ8            1.
    Z2ZZ6ZZ8  NCZZ5ZZ9          -10.
    Z3ZZ6ZZ8  OBJZZZZZ           20.   NCZZ2ZZ8           20.
    Z3ZZ6ZZ8  ACZZ6ZZ8           20.   CCZZ6ZZ8            1.
    Z3ZZ6ZZ8  NCZZ2ZZ9           -5.   ACZZ6ZZ9           -5.
    Z3ZZ6ZZ8  NCZZ5ZZ9          -15.
    Z4ZZ6ZZ8  OBJZZZZZ            1.   NCZZ2ZZ8            1.
    Z4ZZ6ZZ8  ACZZ6ZZ8            1.   NCZZ2ZZ9           -1.
    Z4ZZ6ZZ8  ACZZ6ZZ9           -1.
    Z1ZZ7ZZ8  CCZZ7ZZ8            1.
    Z2ZZ7ZZ8  OBJZZZZZ           10.   NCZZ3ZZ8           10.
    Z2ZZ7ZZ8  ACZZ7ZZ8           10.   CCZZ7ZZ8            1.
    Z2ZZ7ZZ8  NCZZ5ZZ9          -10.
    Z3ZZ7ZZ8  OBJZZZZZ           20.   NCZZ3ZZ8           20.
    Z3ZZ7ZZ8  ACZZ7ZZ8           20.   CCZZ7ZZ8            1.
    Z3ZZ7ZZ8  NCZZ3ZZ9           -5.   ACZZ7ZZ9           -5.
    Z3ZZ7ZZ8  NCZZ5ZZ9          -15.
    Z4ZZ7ZZ8  OBJZZZZZ            1.   NCZZ3ZZ8            1.
    Z4ZZ7ZZ8  ACZZ7ZZ8            1.   NCZZ3ZZ9           -1.
    Z4ZZ7ZZ8  ACZZ7ZZ9           -1.
    Z1ZZ8ZZ8  CCZZ8ZZ8            1.
    Z2ZZ8ZZ8  OBJZZZZZ           30.   NCZZ4ZZ8           30.
    Z2ZZ8ZZ8  ACZZ8ZZ8           30.   CCZZ8ZZ8            1.
    Z2ZZ8ZZ8  NCZZ5ZZ9          -30.
    Z3ZZ8ZZ8  OBJZZZZZ           40.   NCZZ4ZZ8           40.
    Z3ZZ8ZZ8  ACZZ8ZZ8           40.   CCZZ8ZZ8            1.
    Z3ZZ8ZZ8  NCZZ4ZZ9           -7.   ACZZ8ZZ9           -7.
    Z3ZZ8ZZ8  NCZZ5ZZ9          -33.
    Z4ZZ8ZZ8  OBJZZZZZ            1.   NCZZ4ZZ8            1.
    Z4ZZ8ZZ8  ACZZ8ZZ8            1.   NCZZ4ZZ9           -1.
    Z4ZZ8ZZ8  ACZZ8ZZ9           -1.
    Z1ZZ9ZZ8  CCZZ9ZZ8            1.
    Z2ZZ9ZZ8  OBJZZZZZ           30.   NCZZ5ZZ8           30.
    Z2ZZ9ZZ8  ACZZ9ZZ8           30.   CCZZ9ZZ8            1.
    Z2ZZ9ZZ8  NCZZ4ZZ9          -30.
    Z3ZZ9ZZ8  OBJZZZZZ           40.   NCZZ5ZZ8           40.
    Z3ZZ9ZZ8  ACZZ9ZZ8           40.   CCZZ9ZZ8            1.
    Z3ZZ9ZZ8  NCZZ5ZZ9           -7.   ACZZ9ZZ9           -7.
    Z3ZZ9ZZ8  NCZZ4ZZ9          -33.
    Z4ZZ9ZZ8  OBJZZZZZ            1.   NCZZ5ZZ8            1.
    Z4ZZ9ZZ8  ACZZ9ZZ8            1.   NCZZ5ZZ9           -1.
    Z4ZZ9ZZ8  ACZZ9ZZ9           -1.
    Z1Z10ZZ8  CCZ10ZZ8            1.
    Z2Z10ZZ8  OBJZZZZZ           10.   NCZZ4ZZ8           10.
    Z2Z10ZZ8  ACZ10ZZ8           10.   CCZ10ZZ8            1.
    Z2Z10ZZ8  NCZZ6ZZ9          -10.
    Z3Z10ZZ8  OBJZZZZZ           20.   NCZZ4ZZ8           20.
    Z3Z10ZZ8  ACZ10ZZ8           20.   CCZ10ZZ8            1.
    Z3Z10ZZ8  NCZZ4ZZ9           -5.   ACZ10ZZ9           -5.
    Z3Z10ZZ8  NCZZ6ZZ9          -15.
    Z4Z10ZZ8  OBJZZZZZ            1.   NCZZ4ZZ8            1.
    Z4Z10ZZ8  ACZ10ZZ8            1.   NCZZ4ZZ9           -1.
    Z4Z10ZZ8  ACZ10ZZ9           -1.
    Z1Z11ZZ8  CCZ11ZZ8            1.
    Z2Z11ZZ8  OBJZZZZZ           15.   NCZZ6ZZ8           15.
    Z2Z11ZZ8  ACZ11ZZ8           15.   CCZ11ZZ8            1.
    Z3Z11ZZ8  OBJZZZZZ           25.   NCZZ6ZZ8           25.
    Z3Z11ZZ8  ACZ11ZZ8           25.   CCZ11ZZ8            1.
    Z3Z11ZZ8  NCZZ6ZZ9           -7.   ACZ11ZZ9           -7.
    Z4Z11ZZ8  OBJZZZZZ            1.   NCZZ6ZZ8            1.
    Z4Z11ZZ8  ACZ11ZZ8            1.   NCZZ6ZZ9           -1.
    Z4Z11ZZ8  ACZ11ZZ9           -1.
    Z1Z12ZZ8  CCZ12ZZ8            1.
    Z2Z12ZZ8  OBJZZZZZ           50.   NCZZ5ZZ8           50.
    Z2Z12ZZ8  ACZ12ZZ8           50.   CCZ12ZZ8            1.
    Z3Z12ZZ8  OBJZZZZZ           80.   NCZZ5ZZ8           80.
    Z3Z12ZZ8  ACZ12ZZ8           80.   CCZ12ZZ8            1.
    Z3Z12ZZ8  NCZZ5ZZ9          -15.   ACZ12ZZ9          -15.
    Z4Z12ZZ8  OBJZZZZZ            1.   NCZZ5ZZ8            1.
    Z4Z12ZZ8  ACZ12ZZ8            1.   NCZZ5ZZ9           -1.
    Z4Z12ZZ8  ACZ12ZZ9           -1.
    Z1ZZ1ZZ9  CCZZ1ZZ9            1.
    Z2ZZ1ZZ9  OBJZZZZZ           10.   NCZZ1ZZ9           10.
    Z2ZZ1ZZ9  ACZZ1ZZ9           10.   CCZZ1ZZ9            1.
    Z2ZZ1ZZ9  NCZZ2Z10          -10.
    Z3ZZ1ZZ9  OBJZZZZZ           20.   NCZZ1ZZ9           20.
    Z3ZZ1ZZ9  ACZZ1ZZ9           20.   CCZZ1ZZ9            1.
    Z3ZZ1ZZ9  NCZZ1Z10           -5.   ACZZ1Z10           -5.
    Z3ZZ1ZZ9  NCZZ2Z10          -15.
    Z4ZZ1ZZ9  OBJZZZZZ            1.   NCZZ1ZZ9            1.
    Z4ZZ1ZZ9  ACZZ1ZZ9            1.   NCZZ1Z10           -1.
    Z4ZZ1ZZ9  ACZZ1Z10           -1.
    Z1ZZ2ZZ9  CCZZ2ZZ9            1.
    Z2ZZ2ZZ9  OBJZZZZZ           15.   NCZZ2ZZ9           15.
    Z2ZZ2ZZ9  ACZZ2ZZ9           15.   CCZZ2ZZ9            1.
    Z2ZZ2ZZ9  NCZZ3Z10          -15.
    Z3ZZ2ZZ9  OBJZZZZZ           25.   NCZZ2ZZ9           25.
    Z3ZZ2ZZ9  ACZZ2ZZ9           25.   CCZZ2ZZ9            1.
    Z3ZZ2ZZ9  NCZZ2Z10           -7.   ACZZ2Z10           -7.
    Z3ZZ2ZZ9  NCZZ3Z10          -18.
    Z4ZZ2ZZ9  OBJZZZZZ            1.   NCZZ2ZZ9            1.
    Z4ZZ2ZZ9  ACZZ2ZZ9            1.   NCZZ2Z10           -1.
    Z4ZZ2ZZ9  ACZZ2Z10           -1.
    Z1ZZ3ZZ9  CCZZ3ZZ9            1.
    Z2ZZ3ZZ9  OBJZZZZZ           10.   NCZZ2ZZ9           10.
    Z2ZZ3ZZ9  ACZZ3ZZ9           10.   CCZZ3ZZ9            1.
    Z2ZZ3ZZ9  NCZZ1Z10          -10.
    Z3ZZ3ZZ9  OBJZZZZZ           20.   NCZZ2ZZ9           20.
    Z3ZZ3ZZ9  ACZZ3ZZ9           20.   CCZZ3ZZ9            1.
    Z3ZZ3ZZ9  NCZZ2Z10           -5.   ACZZ3Z10           -5.
    Z3ZZ3ZZ9  NCZZ1Z10          -15.
    Z4ZZ3ZZ9  OBJZZZZZ            1.   NCZZ2ZZ9            1.
    Z4ZZ3ZZ9  ACZZ3ZZ9            1.   NCZZ2Z10           -1.
    Z4ZZ3ZZ9  ACZZ3Z10           -1.
    Z1ZZ4ZZ9  CCZZ4ZZ9            1.
    Z2ZZ4ZZ9  OBJZZZZZ           15.   NCZZ1ZZ9           15.
    Z2ZZ4ZZ9  ACZZ4ZZ9           15.   CCZZ4ZZ9            1.
    Z2ZZ4ZZ9  NCZZ4Z10          -15.
    Z3ZZ4ZZ9  OBJZZZZZ           25.   NCZZ1ZZ9           25.
    Z3ZZ4ZZ9  ACZZ4ZZ9           25.   CCZZ4ZZ9            1.
    Z3ZZ4ZZ9  NCZZ1Z10           -7.   ACZZ4Z10           -7.
    Z3ZZ4ZZ9  NCZZ4Z10          -18.
    Z4ZZ4ZZ9  OBJZZZZZ            1.   NCZZ1ZZ9            1.
    Z4ZZ4ZZ9  ACZZ4ZZ9            1.   NCZZ1Z10           -1.
    Z4ZZ4ZZ9  ACZZ4Z10           -1.
    Z1ZZ5ZZ9  CCZZ5ZZ9            1.
    Z2ZZ5ZZ9  OBJZZZZZ           15.   NCZZ2ZZ9           15.
    Z2ZZ5ZZ9  ACZZ5ZZ9           15.   CCZZ5ZZ9            1.
    Z2ZZ5ZZ9  NCZZ4Z10          -15.
    Z3ZZ5ZZ9  OBJZZZZZ           25.   NCZZ2ZZ9           25.
    Z3ZZ5ZZ9  ACZZ5ZZ9           25.   CCZZ5ZZ9            1.
    Z3ZZ5ZZ9  NCZZ2Z10           -7.   ACZZ5Z10           -7.
    Z3ZZ5ZZ9  NCZZ4Z10          -18.
    Z4ZZ5ZZ9  OBJZZZZZ            1.   NCZZ2ZZ9            1.
    Z4ZZ5ZZ9  ACZZ5ZZ9            1.   NCZZ2Z10           -1.
    Z4ZZ5ZZ9  ACZZ5Z10           -1.
    Z1ZZ6ZZ9  CCZZ6ZZ9            1.
    Z2ZZ6ZZ9  OBJZZZZZ           10.   NCZZ2ZZ9           10.
    Z2ZZ6ZZ9  ACZZ6ZZ9           10.   CCZZ6ZZ9            1.
    Z2ZZ6ZZ9  NCZZ5Z10          -10.
    Z3ZZ6ZZ9  OBJZZZZZ           20.   NCZZ2ZZ9           20.
    Z3ZZ6ZZ9  ACZZ6ZZ9           20.   CCZZ6ZZ9            1.
    Z3ZZ6ZZ9  NCZZ2Z10           -5.   ACZZ6Z10           -5.
    Z3ZZ6ZZ9  NCZZ5Z10          -15.
    Z4ZZ6ZZ9  OBJZZZZZ            1.   NCZZ2ZZ9            1.
    Z4ZZ6ZZ9  ACZZ6ZZ9            1.   NCZZ2Z10           -1.
    Z4ZZ6ZZ9  ACZZ6Z10           -1.
    Z1ZZ7ZZ9  CCZZ7ZZ9            1.
    Z2ZZ7ZZ9  OBJZZZZZ           10.   NCZZ3ZZ9           10.
    Z2ZZ7ZZ9  ACZZ7ZZ9           10.   CCZZ7ZZ9            1.
    Z2ZZ7ZZ9  NCZZ5Z10          -10.
    Z3ZZ7ZZ9  OBJZZZZZ           20.   NCZZ3ZZ9           20.
    Z3ZZ7ZZ9  ACZZ7ZZ9           20.   CCZZ7ZZ9            1.
    Z3ZZ7ZZ9  NCZZ3Z10           -5.   ACZZ7Z10           -5.
    Z3ZZ7ZZ9  NCZZ5Z10          -15.
    Z4ZZ7ZZ9  OBJZZZZZ            1.   NCZZ3ZZ9            1.
    Z4ZZ7ZZ9  ACZZ7ZZ9            1.   NCZZ3Z10           -1.
    Z4ZZ7ZZ9  ACZZ7Z10           -1.
    Z1ZZ8ZZ9  CCZZ8ZZ9            1.
    Z2ZZ8ZZ9  OBJZZZZZ           30.   NCZZ4ZZ9           30.
    Z2ZZ8ZZ9  ACZZ8ZZ9           30.   CCZZ8ZZ9            1.
    Z2ZZ8ZZ9  NCZZ5Z10          -30.
    Z3ZZ8ZZ9  OBJZZZZZ           40.   NCZZ4ZZ9           40.
    Z3ZZ8ZZ9  ACZZ8ZZ9           40.   CCZZ8ZZ9            1.
    Z3ZZ8ZZ9  NCZZ4Z10           -7.   ACZZ8Z10           -7.
    Z3ZZ8ZZ9  NCZZ5Z10          -33.
    Z4ZZ8ZZ9  OBJZZZZZ            1.   NCZZ4ZZ9            1.
    Z4ZZ8ZZ9  ACZZ8ZZ9            1.   NCZZ4Z10           -1.
    Z4ZZ8ZZ9  ACZZ8Z10           -1.
    Z1ZZ9ZZ9  CCZZ9ZZ9            1.
    Z2ZZ9ZZ9  OBJZZZZZ           30.   NCZZ5ZZ9           30.
    Z2ZZ9ZZ9  ACZZ9ZZ9           30.   CCZZ9ZZ9            1.
    Z2ZZ9ZZ9  NCZZ4Z10          -30.
    Z3ZZ9ZZ9  OBJZZZZZ           40.   NCZZ5ZZ9           40.
    Z3ZZ9ZZ9  ACZZ9ZZ9           40.   CCZZ9ZZ9            1.
    Z3ZZ9ZZ9  NCZZ5Z10           -7.   ACZZ9Z10           -7.
    Z3ZZ9ZZ9  NCZZ4Z10          -33.
    Z4ZZ9ZZ9  OBJZZZZZ            1.   NCZZ5ZZ9            1.
    Z4ZZ9ZZ9  ACZZ9ZZ9            1.   NCZZ5Z10           -1.
    Z4ZZ9ZZ9  ACZZ9Z10           -1.
    Z1Z10ZZ9  CCZ10ZZ9            1.
    Z2Z10ZZ9  OBJZZZZZ           10.   NCZZ4ZZ9           10.
    Z2Z10ZZ9  ACZ10ZZ9           10.   CCZ10ZZ9            1.
    Z2Z10ZZ9  NCZZ6Z10          -10.
    Z3Z10ZZ9  OBJZZZZZ           20.   NCZZ4ZZ9           20.
    Z3Z10ZZ9  ACZ10ZZ9           20.   CCZ10ZZ9            1.
    Z3Z10ZZ9  NCZZ4Z10           -5.   ACZ10Z10           -5.
    Z3Z10ZZ9  NCZZ6Z10          -15.
    Z4Z10ZZ9  OBJZZZZZ            1.   NCZZ4ZZ9            1.
    Z4Z10ZZ9  ACZ10ZZ9            1.   NCZZ4Z10           -1.
    Z4Z10ZZ9  ACZ10Z10           -1.
    Z1Z11ZZ9  CCZ11ZZ9            1.
    Z2Z11ZZ9  OBJZZZZZ           15.   NCZZ6ZZ9           15.
    Z2Z11ZZ9  ACZ11ZZ9           15.   CCZ11ZZ9            1.
    Z3Z11ZZ9  OBJZZZZZ           25.   NCZZ6ZZ9           25.
    Z3Z11ZZ9  ACZ11ZZ9           25.   CCZ11ZZ9            1.
    Z3Z11ZZ9  NCZZ6Z10           -7.   ACZ11Z10           -7.
    Z4Z11ZZ9  OBJZZZZZ            1.   NCZZ6ZZ9            1.
    Z4Z11ZZ9  ACZ11ZZ9            1.   NCZZ6Z10           -1.
    Z4Z11ZZ9  ACZ11Z10           -1.
    Z1Z12ZZ9  CCZ12ZZ9            1.
    Z2Z12ZZ9  OBJZZZZZ           50.   NCZZ5ZZ9           50.
    Z2Z12ZZ9  ACZ12ZZ9           50.   CCZ12ZZ9            1.
    Z3Z12ZZ9  OBJZZZZZ           80.   NCZZ5ZZ9           80.
    Z3Z12ZZ9  ACZ12ZZ9           80.   CCZ12ZZ9            1.
    Z3Z12ZZ9  NCZZ5Z10          -15.   ACZ12Z10          -15.
    Z4Z12ZZ9  OBJZZZZZ            1.   NCZZ5ZZ9            1.
    Z4Z12ZZ9  ACZ12ZZ9            1.   NCZZ5Z10           -1.
    Z4Z12ZZ9  ACZ12Z10           -1.
    Z1ZZ1Z10  CCZZ1Z10            1.
    Z2ZZ1Z10  OBJZZZZZ           10.   NCZZ1Z10           10.
    Z2ZZ1Z10  ACZZ1Z10           10.   CCZZ1Z10            1.
    Z3ZZ1Z10  OBJZZZZZ           20.   NCZZ1Z10           20.
    Z3ZZ1Z10  ACZZ1Z10           20.   CCZZ1Z10            1.
    Z4ZZ1Z10  OBJZZZZZ            1.   NCZZ1Z10            1.
    Z4ZZ1Z10  ACZZ1Z10            1.
    Z1ZZ2Z10  CCZZ2Z10            1.
    Z2ZZ2Z10  OBJZZZZZ           15.   NCZZ2Z10           15.
    Z2ZZ2Z10  ACZZ2Z10           15.   CCZZ2Z10            1.
    Z3ZZ2Z10  OBJZZZZZ           25.   NCZZ2Z10           25.
    Z3ZZ2Z10  ACZZ2Z10           25.   CCZZ2Z10            1.
    Z4ZZ2Z10  OBJZZZZZ            1.   NCZZ2Z10            1.
    Z4ZZ2Z10  ACZZ2Z10            1.
    Z1ZZ3Z10  CCZZ3Z10            1.
    Z2ZZ3Z10  OBJZZZZZ           10.   NCZZ2Z10           10.
    Z2ZZ3Z10  ACZZ3Z10           10.   CCZZ3Z10            1.
    Z3ZZ3Z10  OBJZZZZZ           20.   NCZZ2Z10           20.
    Z3ZZ3Z10  ACZZ3Z10           20.   CCZZ3Z10            1.
    Z4ZZ3Z10  OBJZZZZZ            1.   NCZZ2Z10            1.
    Z4ZZ3Z10  ACZZ3Z10            1.
    Z1ZZ4Z10  CCZZ4Z10            1.
    Z2ZZ4Z10  OBJZZZZZ           15.   NCZZ1Z10           15.
    Z2ZZ4Z10  ACZZ4Z10           15.   CCZZ4Z10            1.
    Z3ZZ4Z10  OBJZZZZZ           25.   NCZZ1Z10           25.
    Z3ZZ4Z10  ACZZ4Z10           25.   CCZZ4Z10            1.
    Z4ZZ4Z10  OBJZZZZZ            1.   NCZZ1Z10            1.
    Z4ZZ4Z10  ACZZ4Z10            1.
    Z1ZZ5Z10  CCZZ5Z10            1.
    Z2ZZ5Z10  OBJZZZZZ           15.   NCZZ2Z10           15.
    Z2ZZ5Z10  ACZZ5Z10           15.   CCZZ5Z10            1.
    Z3ZZ5Z10  OBJZZZZZ           25.   NCZZ2Z10           25.
    Z3ZZ5Z10  ACZZ5Z10           25.   CCZZ5Z10            1.
    Z4ZZ5Z10  OBJZZZZZ            1.   NCZZ2Z10            1.
    Z4ZZ5Z10  ACZZ5Z10            1.
    Z1ZZ6Z10  CCZZ6Z10            1.
    Z2ZZ6Z10  OBJZZZZZ           10.   NCZZ2Z10           10.
    Z2ZZ6Z10  ACZZ6Z10           10.   CCZZ6Z10            1.
    Z3ZZ6Z10  OBJZZZZZ           20.   NCZZ2Z10           20.
    Z3ZZ6Z10  ACZZ6Z10           20.   CCZZ6Z10            1.
    Z4ZZ6Z10  OBJZZZZZ            1.   NCZZ2Z10            1.
    Z4ZZ6Z10  ACZZ6Z10            1.
    Z1ZZ7Z10  CCZZ7Z10            1.
    Z2ZZ7Z10  OBJZZZZZ           10.   NCZZ3Z10           10.
    Z2ZZ7Z10  ACZZ7Z10           10.   CCZZ7Z10            1.
    Z3ZZ7Z10  OBJZZZZZ           20.   NCZZ3Z10           20.
    Z3ZZ7Z10  ACZZ7Z10           20.   CCZZ7Z10            1.
    Z4ZZ7Z10  OBJZZZZZ            1.   NCZZ3Z10            1.
    Z4ZZ7Z10  ACZZ7Z10            1.
    Z1ZZ8Z10  CCZZ8Z10            1.
    Z2ZZ8Z10  OBJZZZZZ           30.   NCZZ4Z10           30.
    Z2ZZ8Z10  ACZZ8Z10           30.   CCZZ8Z10            1.
    Z3ZZ8Z10  OBJZZZZZ           40.   NCZZ4Z10           40.
    Z3ZZ8Z10  ACZZ8Z10           40.   CCZZ8Z10            1.
    Z4ZZ8Z10  OBJZZZZZ            1.   NCZZ4Z10            1.
    Z4ZZ8Z10  ACZZ8Z10            1.
    Z1ZZ9Z10  CCZZ9Z10            1.
    Z2ZZ9Z10  OBJZZZZZ           30.   NCZZ5Z10           30.
    Z2ZZ9Z10  ACZZ9Z10           30.   CCZZ9Z10            1.
    Z3ZZ9Z10  OBJZZZZZ           40.   NCZZ5Z10           40.
    Z3ZZ9Z10  ACZZ9Z10           40.   CCZZ9Z10            1.
    Z4ZZ9Z10  OBJZZZZZ            1.   NCZZ5Z10            1.
    Z4ZZ9Z10  ACZZ9Z10            1.
    Z1Z10Z10  CCZ10Z10            1.
    Z2Z10Z10  OBJZZZZZ           10.   NCZZ4Z10           10.
    Z2Z10Z10  ACZ10Z10           10.   CCZ10Z10            1.
    Z3Z10Z10  OBJZZZZZ           20.   NCZZ4Z10           20.
    Z3Z10Z10  ACZ10Z10           20.   CCZ10Z10            1.
    Z4Z10Z10  OBJZZZZZ            1.   NCZZ4Z10            1.
    Z4Z10Z10  ACZ10Z10            1.
    Z1Z11Z10  CCZ11Z10            1.
    Z2Z11Z10  OBJZZZZZ           15.   NCZZ6Z10           15.
    Z2Z11Z10  ACZ11Z10           15.   CCZ11Z10            1.
    Z3Z11Z10  OBJZZZZZ           25.   NCZZ6Z10           25.
    Z3Z11Z10  ACZ11Z10           25.   CCZ11Z10            1.
    Z4Z11Z10  OBJZZZZZ            1.   NCZZ6Z10            1.
    Z4Z11Z10  ACZ11Z10            1.
    Z1Z12Z10  CCZ12Z10            1.
    Z2Z12Z10  OBJZZZZZ           50.   NCZZ5Z10           50.
    Z2Z12Z10  ACZ12Z10           50.   CCZ12Z10            1.
    Z3Z12Z10  OBJZZZZZ           80.   NCZZ5Z10           80.
    Z3Z12Z10  ACZ12Z10           80.   CCZ12Z10            1.
    Z4Z12Z10  OBJZZZZZ            1.   NCZZ5Z10            1.
    Z4Z12Z10  ACZ12Z10            1.
RHS
    RHS       NCZZ1ZZ1            5.   NCZZ2ZZ1            5.
    RHS       CCZZ1ZZ1            1.   CCZZ2ZZ1            1.
    RHS       CCZZ3ZZ1            1.   CCZZ4ZZ1            1.
    RHS       CCZZ5ZZ1            1.   CCZZ6ZZ1            1.
    RHS       CCZZ7ZZ1            1.   CCZZ8ZZ1            1.
    RHS       CCZZ9ZZ1            1.   CCZ10ZZ1            1.
    RHS       CCZ11ZZ1            1.   CCZ12ZZ1            1.
    RHS       NCZZ1ZZ2           10.   NCZZ2ZZ2           30.
    RHS       NCZZ3ZZ2           10.   NCZZ4ZZ2            5.
    RHS       CCZZ1ZZ2            1.   CCZZ2ZZ2            1.
    RHS       CCZZ3ZZ2            1.   CCZZ4ZZ2            1.
    RHS       CCZZ5ZZ2            1.   CCZZ6ZZ2            1.
    RHS       CCZZ7ZZ2            1.   CCZZ8ZZ2            1.
    RHS       CCZZ9ZZ2            1.   CCZ10ZZ2            1.
    RHS       CCZ11ZZ2            1.   CCZ12ZZ2            1.
    RHS       NCZZ1ZZ3           20.   NCZZ2ZZ3           40.
    RHS       NCZZ3ZZ3           25.   NCZZ4ZZ3           10.
    RHS       NCZZ5ZZ3           15.   CCZZ1ZZ3            1.
    RHS       CCZZ2ZZ3            1.   CCZZ3ZZ3            1.
    RHS       CCZZ4ZZ3            1.   CCZZ5ZZ3            1.
    RHS       CCZZ6ZZ3            1.   CCZZ7ZZ3            1.
    RHS       CCZZ8ZZ3            1.   CCZZ9ZZ3            1.
    RHS       CCZ10ZZ3            1.   CCZ11ZZ3            1.
    RHS       CCZ12ZZ3            1.   NCZZ1ZZ4           10.
    RHS       NCZZ2ZZ4           30.   NCZZ3ZZ4           20.
    RHS       NCZZ4ZZ4           20.   NCZZ5ZZ4           40.
    RHS       CCZZ1ZZ4            1.   CCZZ2ZZ4            1.
    RHS       CCZZ3ZZ4            1.   CCZZ4ZZ4            1.
    RHS       CCZZ5ZZ4            1.   CCZZ6ZZ4            1.
    RHS       CCZZ7ZZ4            1.   CCZZ8ZZ4            1.
    RHS       CCZZ9ZZ4            1.   CCZ10ZZ4            1.
    RHS       CCZ11ZZ4            1.   CCZ12ZZ4            1.
    RHS       NCZZ1ZZ5            3.   NCZZ2ZZ5           25.
    RHS       NCZZ3ZZ5           15.   NCZZ4ZZ5           15.
    RHS       NCZZ5ZZ5           30.   NCZZ6ZZ5            5.
    RHS       CCZZ1ZZ5            1.   CCZZ2ZZ5            1.
    RHS       CCZZ3ZZ5            1.   CCZZ4ZZ5            1.
    RHS       CCZZ5ZZ5            1.   CCZZ6ZZ5            1.
    RHS       CCZZ7ZZ5            1.   CCZZ8ZZ5            1.
    RHS       CCZZ9ZZ5            1.   CCZ10ZZ5            1.
    RHS       CCZ11ZZ5            1.   CCZ12ZZ5            1.
    RHS       NCZZ2ZZ6           20.   NCZZ3ZZ6           10.
    RHS       NCZZ4ZZ6           10.   NCZZ5ZZ6           20.
    RHS       NCZZ6ZZ6           15.   CCZZ1ZZ6            1.
    RHS       CCZZ2ZZ6            1.   CCZZ3ZZ6            1.
    RHS       CCZZ4ZZ6            1.   CCZZ5ZZ6            1.
    RHS       CCZZ6ZZ6            1.   CCZZ7ZZ6            1.
    RHS       CCZZ8ZZ6            1.   CCZZ9ZZ6            1.
    RHS       CCZ10ZZ6            1.   CCZ11ZZ6            1.
    RHS       CCZ12ZZ6            1.   NCZZ2ZZ7           10.
    RHS       NCZZ4ZZ7           10.   NCZZ5ZZ7           10.
    RHS       NCZZ6ZZ7           20.   CCZZ1ZZ7            1.
    RHS       CCZZ2ZZ7            1.   CCZZ3ZZ7            1.
    RHS       CCZZ4ZZ7            1.   CCZZ5ZZ7            1.
    RHS       CCZZ6ZZ7            1.   CCZZ7ZZ7            1.
    RHS       CCZZ8ZZ7            1.   CCZZ9ZZ7            1.
    RHS       CCZ10ZZ7            1.   CCZ11ZZ7            1.
    RHS       CCZ12ZZ7            1.   NCZZ4ZZ8            5.
    RHS       NCZZ6ZZ8           10.   CCZZ1ZZ8            1.
    RHS       CCZZ2ZZ8            1.   CCZZ3ZZ8            1.
    RHS       CCZZ4ZZ8            1.   CCZZ5ZZ8            1.
    RHS       CCZZ6ZZ8            1.   CCZZ7ZZ8            1.
    RHS       CCZZ8ZZ8            1.   CCZZ9ZZ8            1.
    RHS       CCZ10ZZ8            1.   CCZ11ZZ8            1.
    RHS       CCZ12ZZ8            1.   NCZZ6ZZ9            5.
    RHS       CCZZ1ZZ9            1.   CCZZ2ZZ9            1.
    RHS       CCZZ3ZZ9            1.   CCZZ4ZZ9            1.
    RHS       CCZZ5ZZ9            1.   CCZZ6ZZ9            1.
    RHS       CCZZ7ZZ9            1.   CCZZ8ZZ9            1.
    RHS       CCZZ9ZZ9            1.   CCZ10ZZ9            1.
    RHS       CCZ11ZZ9            1.   CCZ12ZZ9            1.
    RHS       CCZZ1Z10            1.   CCZZ2Z10            1.
    RHS       CCZZ3Z10            1.   CCZZ4Z10            1.
    RHS       CCZZ5Z10            1.   CCZZ6Z10            1.
    RHS       CCZZ7Z10            1.   CCZZ8Z10            1.
    RHS       CCZZ9Z10            1.   CCZ10Z10            1.
    RHS       CCZ11Z10            1.   CCZ12Z10            1.
ENDATA
