"""Seeded, platform-independent random streams for synthetic CAMP inputs.

splitmix64 seeds xoshiro256** (Blackman & Vigna, "Scrambled linear
pseudorandom number generators", ACM TOMS 47(4), 2021).  Two front ends
produce the same numbers: a scalar one (pure Python integers) for the
mechanism structure and a vectorised one (numpy uint64) that runs one
independent stream per grid cell, so a cell's inputs do not depend on how
many cells are generated alongside it.
"""
import numpy as np

_MASK = (1 << 64) - 1
_GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(x):
    """One splitmix64 step: returns (new_state, output)."""
    x = (x + _GOLDEN) & _MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & _MASK


class Xoshiro:
    """Scalar xoshiro256** stream."""

    def __init__(self, seed):
        st = seed & _MASK
        s = []
        for _ in range(4):
            st, out = splitmix64(st)
            s.append(out)
        self.s = s

    def next_u64(self):
        s0, s1, s2, s3 = self.s
        result = (_rotl((s1 * 5) & _MASK, 7) * 9) & _MASK
        t = (s1 << 17) & _MASK
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = _rotl(s3, 45)
        self.s = [s0, s1, s2, s3]
        return result

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def log_uniform(self, lo, hi):
        return float(np.exp(self.uniform(np.log(lo), np.log(hi))))

    def randint(self, lo, hi):
        """Integer uniform on [lo, hi] (inclusive)."""
        n = hi - lo + 1
        return lo + int(self.uniform() * n) % n

    def sample(self, n, k):
        """k distinct integers from range(n) (partial Fisher-Yates)."""
        pool = list(range(n))
        for i in range(k):
            j = i + int(self.uniform() * (n - i)) % (n - i)
            pool[i], pool[j] = pool[j], pool[i]
        return pool[:k]


class CellStreams:
    """One xoshiro256** stream per cell, advanced in lock step (numpy)."""

    def __init__(self, seed, cell_ids):
        cell_ids = np.asarray(cell_ids, dtype=np.uint64)
        st = (np.uint64(seed & _MASK) * np.uint64(0x100000001B3)) ^ cell_ids
        s = []
        with np.errstate(over="ignore"):
            for _ in range(4):
                st = st + np.uint64(_GOLDEN)
                z = st.copy()
                z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
                z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
                s.append(z ^ (z >> np.uint64(31)))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return (x << np.uint64(k)) | (x >> np.uint64(64 - k))

    def next_u64(self):
        s0, s1, s2, s3 = self.s
        with np.errstate(over="ignore"):
            result = self._rotl(s1 * np.uint64(5), 7) * np.uint64(9)
        t = s1 << np.uint64(17)
        s2 = s2 ^ s0
        s3 = s3 ^ s1
        s1 = s1 ^ s2
        s0 = s0 ^ s3
        s2 = s2 ^ t
        s3 = self._rotl(s3, 45)
        self.s = [s0, s1, s2, s3]
        return result

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u64() >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def log_uniform(self, lo, hi):
        return np.exp(self.uniform(np.log(lo), np.log(hi)))
