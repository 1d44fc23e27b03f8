"""CPU oracle of the epcq path: test infrastructure only (the checker)."""
