"""Single-GPU stand-in for parallel/mpi_utils.py:12-48.

One process drives one GPU and owns the whole domain: dims (1, 1), coords
(0, 0), every neighbour absent, so every side of the array is a physical edge
(solvers/roe_exner_solver.py:27-35) and halo exchange is the identity
(roe_exner.py:782-783).  parNx * parNy > 1 is rejected.
"""


class Parallel:
    def __init__(self, params=None):
        params = params or {}
        self.parNx = int(params.get("parNx", 1))
        self.parNy = int(params.get("parNy", 1))
        if self.parNx * self.parNy != 1:
            raise NotImplementedError("this backend runs one domain per GPU (parNx = parNy = 1)")
        self.rank, self.size = 0, 1
        self.dims = (1, 1)
        self.coords = (0, 0)
        self.periods = (False, False)
        self.neighbors = {"north": None, "south": None, "east": None, "west": None}
        self.cart_comm = None


def identity_halo(arr):
    return arr
