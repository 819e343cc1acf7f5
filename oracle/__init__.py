"""CPU oracle package -- test infrastructure only (tests/, smoke(), bench cpu_baseline)."""
