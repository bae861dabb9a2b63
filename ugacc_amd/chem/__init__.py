"""Host chemistry front end: basis sets, Gaussian integrals, RHF, the MO
Hamiltonian blocks and the seeded water-cluster geometries."""
