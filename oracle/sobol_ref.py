"""Owen-scrambled Sobol numbers in numpy: TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the bench's CPU leg import this module, to
check the device's numbers; the product never does.

Restates the reference's sequence member (plugins/sampler/sobol/sobol.cpp):
28 coordinates (the pixel, the lens and three events of 8), each its own
Sobol dimension with Joe and Kuo's direction numbers over 16 bits; the 6
shared coordinates take dimensions 0..5 and the later ones take dimensions
0.. again on a second index shuffle (sobol.h:13-21, sobol.cpp:129-136); the
sample's place shuffled by a Laine-Karras hash of its reversed 16 bits
(sobol.cpp:40-42) and every value Owen-scrambled the same way, seeded by the
upper half of the stream (sobol.cpp:138-145).

Pinned by tests/test_ref_pins.py against the reference's own functions
(oracle/_ref, built from the reference's sobol.cpp) and fixtures written from
them, against scipy's independent Sobol generator (the unscrambled points
under the Gray-code order it draws them in), and by the reference's own net
and evenness checks (target/tests/sampler.cpp:378-435).
"""
import numpy as np

SOBOL_DIMENSIONS = 28  # sobol.h:13
SHARED_COORDINATES = 6  # sobol.h:21
SOBOL_BITS = 16  # sobol.cpp:44

# (degree, polynomial, first direction numbers) for each dimension: Joe and
# Kuo's table (new-joe-kuo-6.21201), the reference's sobolRules
# (sobol.cpp:55-84); dimension 0 is van der Corput.
RULES = (
    (0, 0, (1,)), (1, 0, (1,)), (2, 1, (1, 3)), (3, 1, (1, 3, 1)), (3, 2, (1, 1, 1)),
    (4, 1, (1, 1, 3, 3)), (4, 4, (1, 3, 5, 13)), (5, 2, (1, 1, 5, 5, 17)), (5, 4, (1, 1, 5, 5, 5)),
    (5, 7, (1, 1, 7, 11, 19)), (5, 11, (1, 1, 5, 1, 1)), (5, 13, (1, 1, 1, 3, 11)), (5, 14, (1, 3, 5, 5, 31)),
    (6, 1, (1, 3, 3, 9, 7, 49)), (6, 13, (1, 1, 1, 15, 21, 21)), (6, 16, (1, 3, 1, 13, 27, 49)),
    (6, 19, (1, 1, 1, 15, 7, 5)), (6, 22, (1, 3, 1, 15, 13, 25)), (6, 25, (1, 1, 5, 5, 19, 61)),
    (7, 1, (1, 3, 7, 11, 23, 15, 103)), (7, 4, (1, 3, 7, 13, 13, 15, 69)), (7, 7, (1, 1, 3, 13, 7, 35, 63)),
    (7, 8, (1, 3, 5, 9, 1, 25, 53)), (7, 14, (1, 3, 1, 13, 9, 35, 107)), (7, 19, (1, 3, 1, 5, 27, 61, 31)),
    (7, 21, (1, 1, 5, 11, 19, 41, 61)), (7, 28, (1, 3, 5, 3, 3, 13, 69)), (7, 31, (1, 1, 7, 13, 1, 19, 1)),
)

M32 = np.uint64(0xFFFFFFFF)


def _u32(v):
    return np.asarray(v, dtype=np.uint64) & M32


def stirred(v):
    """The 32-bit finaliser (sobol.cpp:8-14)."""
    v = _u32(v)
    v = v ^ (v >> np.uint64(16))
    v = (v * np.uint64(0x7FEB352D)) & M32
    v = v ^ (v >> np.uint64(15))
    v = (v * np.uint64(0x846CA68B)) & M32
    return v ^ (v >> np.uint64(16))


def reversed_bits(v):
    """The 32 bits in reverse order (sobol.cpp:16-22)."""
    v = _u32(v)
    out = np.zeros_like(v)
    for b in range(32):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(31 - b)
    return out


def reversed_sixteen(v):
    """The low 16 bits in reverse order (sobol.cpp:24-29)."""
    v = _u32(v)
    out = np.zeros_like(v)
    for b in range(16):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(15 - b)
    return out


def laine_karras(v, seed):
    """Laine and Karras's nested scramble over the low bits (sobol.cpp:31-37):
    every output bit depends on the input bits at and below it."""
    v = (_u32(v) + _u32(seed)) & M32
    for c in (0x6C50B47C, 0xB82F1E52, 0xC7AFE638):
        v = v ^ ((v * np.uint64(c)) & M32)
    return v ^ ((v * np.uint64(0x8D22F6E6)) & M32)


def shuffled_index(index, seed):
    """The sample's place on the sequence (sobol.cpp:39-41)."""
    return reversed_sixteen(laine_karras(reversed_sixteen(_u32(index) & np.uint64(0xFFFF)), seed) & np.uint64(0xFFFF))


def direction_numbers(rule):
    """The 16 direction numbers m_1.. of a dimension: the given ones, then
    the recurrence of its primitive polynomial (sobol.cpp:86-103)."""
    degree, poly, first = rule
    m = []
    for bit in range(SOBOL_BITS):
        if degree == 0:
            m.append(first[0])
        elif bit < degree:
            m.append(first[bit])
        else:
            n = m[bit - degree] ^ (m[bit - degree] << degree)
            for k in range(1, degree):
                if (poly >> (degree - 1 - k)) & 1:
                    n ^= m[bit - k] << k
            m.append(n)
    return m


def _columns():
    """Per dimension, the 16 bit-reversed columns placedBits XORs
    (sobol.cpp:105-112): column b holds m_b << (31 - b) reversed, so that
    the value's bit 0 is the fraction's most significant bit."""
    cols = np.zeros((SOBOL_DIMENSIONS, SOBOL_BITS), dtype=np.uint64)
    for d, rule in enumerate(RULES):
        m = direction_numbers(rule)
        for b in range(SOBOL_BITS):
            cols[d, b] = reversed_bits((m[b] << (31 - b)) & 0xFFFFFFFF)
    return cols


COLUMNS = _columns()


def placed_bits(dimension, place):
    """The unscrambled Sobol value of a place, bit-reversed (sobol.cpp:105-112)."""
    place = _u32(place)
    value = np.zeros(np.broadcast(np.asarray(dimension), place).shape, dtype=np.uint64)
    dimension = np.broadcast_to(np.asarray(dimension, dtype=np.int64), value.shape)
    for b in range(SOBOL_BITS):
        on = ((place >> np.uint64(b)) & np.uint64(1)) != 0
        value ^= np.where(on, COLUMNS[dimension, b], np.uint64(0))
    return value


def unscrambled(dimension, index):
    """Point `index` of Sobol dimension `dimension` as a fraction, no
    scramble and no shuffle."""
    return reversed_bits(placed_bits(dimension, index)).astype(np.float64) * 2.0 ** -32


def sobol_number(stream, index, coordinate):
    """sobolNumber (sobol.cpp:146-148) for a coordinate below 28, in double."""
    stream = np.asarray(stream, dtype=np.uint64)
    coordinate = np.asarray(coordinate, dtype=np.uint64)
    scramble = stream >> np.uint64(32)
    own = np.where(coordinate < np.uint64(SHARED_COORDINATES), np.uint64(0), np.uint64(0x68E31DA4))
    place = shuffled_index(index, stirred((scramble + own) & M32))
    seed = stirred((scramble + ((coordinate * np.uint64(0xA136AAAD)) & M32) + np.uint64(0x2545F491)) & M32)
    dim = np.where(coordinate < np.uint64(SHARED_COORDINATES), coordinate, coordinate - np.uint64(SHARED_COORDINATES))
    value = placed_bits(dim.astype(np.int64), place)
    return reversed_bits(laine_karras(value, seed)).astype(np.float64) * 2.0 ** -32
