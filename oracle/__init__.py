"""FP64 CPU oracle of the TLS search -- test infrastructure (tests, smoke, CPU baseline)."""
