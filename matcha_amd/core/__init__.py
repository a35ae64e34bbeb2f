"""MI355X implementations of the reference's hot-path classes (same names)."""
