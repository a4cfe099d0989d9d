"""Named benchmark workloads (synthetic integrals from a seed, see jci_synth_integrals).

fci12: FCI 4a4b in 12 orbitals (245,025 determinants), the reference's CPU-runnable case.
"""
WORKLOADS = {
    "fci12": dict(na=4, nb=4, norbs=12, ndocc=0, nactv=12, xlvl=0, rule=1, seed=100, nroots=1,
                  desc="FCI 8e/12o (4a4b), 245,025 determinants, 1 root, residual 1e-6"),
}
