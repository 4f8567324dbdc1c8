"""std::mt19937_64 (Matsumoto & Nishimura's 64-bit Mersenne Twister with the
parameters the C++ standard fixes, [rand.predef]) for replaying libctc's
seeded ALS mode rotation in tests.  Pinned by the standard's own check value:
the 10000th output of a default-seeded (5489) engine is 9981545732273789042."""

_W, _N, _M, _R = 64, 312, 156, 31
_A = 0xB5026F5AA96619E9
_U, _D = 29, 0x5555555555555555
_S, _B = 17, 0x71D67FFFEDA60000
_T, _C = 37, 0xFFF7EEE000000000
_L = 43
_F = 6364136223846793005
_MASK = (1 << 64) - 1
_LOWER = (1 << _R) - 1
_UPPER = _MASK & ~_LOWER


class MT19937_64:
    def __init__(self, seed=5489):
        self.mt = [0] * _N
        self.mt[0] = seed & _MASK
        for i in range(1, _N):
            self.mt[i] = (_F * (self.mt[i - 1] ^ (self.mt[i - 1] >> (_W - 2))) + i) & _MASK
        self.idx = _N

    def _twist(self):
        mt = self.mt
        for i in range(_N):
            x = (mt[i] & _UPPER) | (mt[(i + 1) % _N] & _LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= _A
            mt[i] = mt[(i + _M) % _N] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= _N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> _U) & _D
        y ^= (y << _S) & _B
        y ^= (y << _T) & _C
        y ^= y >> _L
        return y & _MASK
