"""CPU checkers for the tests, the smoke check and bench.py's CPU baseline (never the product)."""
