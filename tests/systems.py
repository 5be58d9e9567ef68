"""Molecules, bases and densities shared by the tests."""
import numpy as np

from gradscf_amd import molecule as M
from gradscf_amd.basis import prepare_basis
from gradscf_amd.packing import pack


def tables(spec, basis, cart=True):
    topo, par = prepare_basis(spec, basis, cart=cart)
    atm, bas, env = pack(topo, par)
    return topo, par, atm, bas, env


def water():
    return M.water_monomer()


def water_dimer():
    return M.water_cluster(2, seed=7, box=4.0)


def random_density(n, nb=1, seed=0, symmetric=True):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(nb, n, n))
    if symmetric:
        d = 0.5 * (d + d.transpose(0, 2, 1))
    return d
