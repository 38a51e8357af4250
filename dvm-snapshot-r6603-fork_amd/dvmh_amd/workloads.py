"""The converted workload programs as Python objects: setup / run / teardown
through their C entry points, plus the source positions (file, line) of the
loops their iteration phase executes, which the tuner keys its sites by."""
import ctypes

from . import DEVICE_TYPE_CUDA, DEVICE_TYPE_HOST, REGION_AMD_GRAPH, SEED_BASE, Program

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


class Workload:
    name = ''
    cdv = b''
    lines = ()
    seed_offset = 0
    lazy = False            # regions deferred (lazy, never captured while swept)
    record_prefixes = None  # tuning records a sweep of it replaces (default: its program's)

    def __init__(self, size, targets=DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA, flags=0):
        self.size = size
        self.P = Program(self.name)
        self.P.set_targets(targets, flags)

    def setup(self):
        raise NotImplementedError

    def run(self, iters):
        raise NotImplementedError

    def teardown(self):
        self.P.fn('teardown', [])()

    def points_per_iter(self):
        """Point-updates per iteration: interior points written by the loops
        that write the grid (a point written by two loops of one iteration,
        as in 2-D Jacobi's stencil and copy loops, counts once)."""
        return (self.size - 2) ** 3

    def bytes_per_iter(self):
        """Algorithmic bytes per iteration: every loop reads each array it uses
        once and writes each array it writes once, 8 B per FP64 element."""
        return 16 * self.points_per_iter()


class Jac2d(Workload):
    name, cdv, lines, seed_offset = 'jac2d', b'jac2d.cdv', (38, 43), 1

    def setup(self):
        self.P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(self.size, SEED_BASE + self.seed_offset)

    def run(self, iters):
        return self.P.fn('iterate', [ctypes.c_int, _dp], ctypes.c_double)(iters, None)

    def points_per_iter(self):
        return (self.size - 2) ** 2

    def bytes_per_iter(self):
        # stencil loop: read A, write B (16 B); copy loop with MAX: read A and B, write A (24 B)
        return 40 * self.points_per_iter()

    def fused_bytes_per_iter(self):
        # the two loops as one fused pass: read A, write B and A (24 B)
        return 24 * self.points_per_iter()


class Jac3d(Workload):
    name, cdv, lines, seed_offset = 'jac3d', b'jac3d.cdv', (41, 54), 2

    def setup(self):
        self.P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(self.size, SEED_BASE + self.seed_offset)

    def run(self, iters):
        return self.P.fn('iterate', [ctypes.c_int, _dp, _dp], ctypes.c_double)(iters, None, None)


class Sor3d(Workload):
    name, cdv, lines, seed_offset = 'sor3d', b'sor3d.cdv', (32,), 3

    def setup(self):
        self.P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(self.size, 1.5, SEED_BASE + self.seed_offset)

    def run(self, iters):
        return self.P.fn('iterate', [ctypes.c_int, _dp], ctypes.c_double)(iters, None)


class Sor3dChain(Sor3d):
    """The SOR sweeps `group` at a time in one batch (sor3d_iterate_chain):
    each group runs through the sweep loop's chain handler."""
    group = 30

    def run(self, iters):
        return self.P.fn('iterate_chain', [ctypes.c_int, ctypes.c_int, _dp], ctypes.c_double)(iters, self.group, None)


class Adi3d(Workload):
    name, cdv, lines, seed_offset = 'adi3d', b'adi3d.cdv', (34, 39, 44, 49, 54, 59, 64), 4

    def setup(self):
        self.P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(self.size, 0.3, SEED_BASE + self.seed_offset)

    def run(self, iters):
        return self.P.fn('iterate', [ctypes.c_int, _dp, _lp], ctypes.c_double)(iters, None, None)

    def points_per_iter(self):
        # six line sweeps, each updating (n - 1) points of every line of n^2 lines
        return 6 * (self.size - 1) * self.size ** 2

    def bytes_per_iter(self):
        # each sweep reads and writes its points (16 B per point-update); the
        # MAXLOC loop reads the whole array once (8 B per element)
        return 16 * self.points_per_iter() + 8 * self.size ** 3


class Jac2dPairs(Jac2d):
    """2-D Jacobi with its region deferred to the region's end (the program's
    graph flag), so the stencil loop and the copy + MAX loop run as one fused
    pair: the tuner times the pair site (first loop's line, pair kind)."""
    name = 'jac2d'
    lines = (38,)
    lazy = True
    record_prefixes = ('jac2d.cdv:38|k6|',)

    def __init__(self, size, targets=DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA, flags=0):
        super().__init__(size, targets, flags | REGION_AMD_GRAPH)


ALL = {'jac2d': Jac2d, 'jac2d_pairs': Jac2dPairs, 'jac3d': Jac3d, 'sor3d': Sor3d, 'sor3d_chain': Sor3dChain, 'adi3d': Adi3d}
