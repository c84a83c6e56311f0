"""Seeded workloads shared by the tests (small enough for the oracle)."""
import functools

from exciting_amd import synth


@functools.lru_cache(maxsize=None)
def system(name: str):
    if name == "tiny":
        return synth.build(synth.tiny(seed=7))
    if name == "tiny_lmaxmat":
        return synth.build(synth.tiny(seed=8, lmax=4, lmaxmat=2, natoms=(2, 1)))
    if name == "tiny_p3":
        return synth.build(synth.tiny(seed=9, lmax=3, apword=3, nlo_l=(0, 1, 1)))
    if name == "tiny_nolo":
        return synth.build(synth.tiny(seed=10, nlo_l=()))
    if name == "tiny_rel":
        return synth.add_relativity(synth.build(synth.tiny(seed=11)), seed=3, iora=True)
    if name == "tiny_zora":
        return synth.add_relativity(synth.build(synth.tiny(seed=12)), seed=4, iora=False)
    if name == "tiny_multi":
        return synth.build(synth.tiny(seed=13, nk=5))
    if name == "tiny_lmaxmat_multi":
        return synth.build(synth.tiny(seed=14, lmax=4, lmaxmat=2, natoms=(2, 1), nk=3))
    if name == "tiny_rel_multi":
        return synth.add_relativity(synth.build(synth.tiny(seed=15, nk=3)), seed=6, iora=True)
    if name == "sc8_4k":
        return synth.build(synth.config(1), kpoints=[0, 3, 5, 6])
    if name == "ortho16_2k":
        return synth.build(synth.config(2), kpoints=[0, 77])
    if name == "fcc2":
        return synth.build(synth.config(0))
    if name == "sc8_k0":
        return synth.build(synth.config(1), kpoints=[0, 5])
    if name == "sc64_k0":
        return synth.build(synth.config(3), kpoints=[0])
    raise KeyError(name)
