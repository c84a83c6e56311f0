"""exciting_amd: MI355X-native LAPW Hamiltonian/overlap setup for exciting."""
