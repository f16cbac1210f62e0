"""Small decks shared by the tests (generated into a temporary directory)."""
import pathlib

import tild_gen

GAUSS3 = [("gaussian", 1, 1, 1.0, 0.5), ("gaussian", 2, 2, 1.0, 0.5), ("gaussian", 1, 2, 2.0, 0.5)]
MIXED = GAUSS3 + [("erf", 1, 2, 0.1, 1.0, 0.5), ("gaussian_erf", 2, 2, 0.5, 0.6, 0.5)]


def melt3d(tmp, n_chains=40, N=10, box=(6.0, 6.0, 6.0), grid=(12, 12, 12), styles=None, integrator="EM",
           delt=0.005, seed=3, pmeorder=1):
    d = pathlib.Path(tmp)
    text = tild_gen.melt(d, n_chains, N, box, grid, styles=styles or GAUSS3, integrator=integrator,
                         delt=delt, seed=seed, pmeorder=pmeorder)
    return text, d


def melt2d(tmp, n_chains=60, N=10, box=(10.0, 8.0), grid=(20, 16), styles=None, integrator="EM",
           delt=0.005, seed=4, pmeorder=1):
    d = pathlib.Path(tmp)
    text = tild_gen.melt(d, n_chains, N, (box[0], box[1], 1.0), grid, dim=2, styles=styles or GAUSS3,
                         integrator=integrator, delt=delt, seed=seed, pmeorder=pmeorder)
    return text, d


def charged(tmp, seed=5, grid=(16, 16, 18)):
    d = pathlib.Path(tmp)
    text = tild_gen.coacervate(d, n_chains_each=12, N=8, n_counter_pairs=20, total=600, box=(7.0, 7.0, 8.0),
                               grid=grid, seed=seed)
    return text, d
