"""Mirror of the reference package level of the same name."""
