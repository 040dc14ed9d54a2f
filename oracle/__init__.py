"""CPU oracle of the parry3d hot path -- test infrastructure only (see oracle.py)."""
