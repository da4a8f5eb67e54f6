"""CPU oracle of the optical-flow path (test infrastructure only)."""
