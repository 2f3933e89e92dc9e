"""PBRT v4 (b0bdf461) Sobol known answers, as data.

Restated from the reference's tests/sampling_tests.cpp:26-135 (literal
values generated there with PBRT in Float mode): interval index, MurmurHash
dimension hash, the word for each randomization (none, permutedigits,
fastowen, owen) and the binary32 camera remaps (hex floats).
"""
KATS = [
    dict(m=0, x=0, y=0, sample=0, dimension=3, seed=0, index=0, hash=0x6E97A272,
         words=[0x00000000, 0x6E97A272, 0x410C32C4, 0x4D50AAAC], camera=["0x0p+0", "0x0p+0"]),
    dict(m=0, x=0, y=0, sample=4503599627370495, dimension=1023, seed=-2147483648, index=4503599627370495,
         hash=0xAE830AFD, words=[0x0C76950B, 0xA2F59FF6, 0x95CFA4E2, 0xD2C575A4],
         camera=["0x1.fffffep-1", "0x1.0011p-4"]),
    dict(m=1, x=1, y=0, sample=15, dimension=11, seed=-1, index=63, hash=0x524AFD16,
         words=[0x24000000, 0x764AFD16, 0x5851B677, 0x7FA0FFCD], camera=["0x1.fp-1", "0x1.5p-1"]),
    dict(m=3, x=4, y=1, sample=9, dimension=46, seed=17, index=609, hash=0x60D98904,
         words=[0xCD400000, 0xAD998904, 0xCAE056C4, 0xC1BE889E], camera=["0x1.9p-3", "0x1.bp-3"]),
    dict(m=6, x=35, y=0, sample=111, dimension=16, seed=-19, index=456689, hash=0x9A34955C,
         words=[0x40D1A000, 0xDAE5355C, 0x7DD112E0, 0x0ECC6E63], camera=["0x1.f7bp-1", "0x1.e2cp-3"]),
    dict(m=12, x=2244, y=33, sample=65535, dimension=7, seed=889, index=1099506831921, hash=0x70C9554B,
         words=[0x6E2AEDBD, 0x1EE3B8F6, 0xF3FB074A, 0xFAAE0197], camera=["0x1.6dcp-1", "0x1.9b4p-1"]),
    dict(m=19, x=254427, y=4, sample=16383, dimension=984, seed=2147483647, index=4503458496306238,
         hash=0xD776BC96, words=[0xABAA0426, 0x7CDCB8B0, 0xB09E7B61, 0xC1EFCB3F], camera=["0x1.8p-1", "0x1.58cp-1"]),
    dict(m=25, x=33554431, y=33554430, sample=3, dimension=1021, seed=-79, index=3956104878882815,
         hash=0x21AEC8AC, words=[0x5C681095, 0x7DC6D839, 0x5A6FC5D2, 0x300114FD], camera=["0x0p+0", "0x0p+0"]),
]
