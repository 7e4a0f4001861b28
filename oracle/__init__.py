"""TEST INFRASTRUCTURE: the CPU oracle of the fluid step (see oracle.py)."""
