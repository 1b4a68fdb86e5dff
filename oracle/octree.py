"""Pure-Python restatement of the octree index codec (test infrastructure only).

cuda/src/utils.cu:129-189: octpath = sum over levels l of (i_l << 2 | j_l << 1 | k_l) << 3 l, then
shifted left by 3 (16 - level) so the path is left-aligned in 48 bits; the inverse undoes it.

Pinning (tests/test_oracle_misc.py): hand-computed values (e.g. level 1 child 5 -> 5 << 45),
round trips, and the property that the left-aligned path order equals depth-first (Morton)
order of the cells.
"""
MAX_NUM_LEVELS = 16


def ijk_to_octpath(i: int, j: int, k: int, level: int) -> int:
    path = 0
    for l in range(level):
        bits = (((i >> l) & 1) << 2) | (((j >> l) & 1) << 1) | ((k >> l) & 1)
        path |= bits << (3 * l)
    return path << (3 * (MAX_NUM_LEVELS - level))


def octpath_to_ijk(path: int, level: int):
    path >>= 3 * (MAX_NUM_LEVELS - level)
    i = j = k = 0
    for l in range(level):
        b = path & 7
        i |= ((b >> 2) & 1) << l
        j |= ((b >> 1) & 1) << l
        k |= (b & 1) << l
        path >>= 3
    return i, j, k
