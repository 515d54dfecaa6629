"""CPU oracle of the V1 path: test infrastructure only (tests/, the smoke check and bench.py's CPU baseline).
The product never imports it."""
