"""ctypes mirrors of the COMPAS structures in include/emstr_us.h and
include/emstr_mesh.h (field order and types must match the header)."""
import ctypes as C

import numpy as np

_I = C.c_int
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PPI = C.POINTER(_PI)
_PD = C.POINTER(C.c_double)
_PPD = C.POINTER(_PD)

ORDER1 = 0x000002
ORDER2 = 0x000004
QUICKEST = 0x000020
VANLEER = 0x000080
FFSL = 0x080000
NONE = 0x80000000
MONOTONIC = 0x0010
CLIP = 0x8000000
V6898 = 0x000800
F_FFSL = 0x100
SUB_STEP = 2


class Geometry(C.Structure):
    _fields_ = [
        ("szc", _I), ("sze", _I), ("szcS", _I), ("szeS", _I),
        ("nobc", _I), ("nz", _I), ("npem", _I), ("wn", _I),
        ("layers", _PD), ("s2i", _PI), ("s2k", _PI),
        ("npe", _PI), ("c2c", _PPI), ("c2e", _PPI), ("e2c", _PPI), ("e2e", _PPI),
        ("eSc", _PPI), ("ep", _PI), ("em", _PI), ("m2d", _PI), ("m2de", _PI),
        ("zp1", _PI), ("zm1", _PI), ("zp1e", _PI), ("zm1e", _PI),
        ("w3_t", _PI), ("w3_e1", _PI),
        ("n3_t", _I), ("v3_t", _I), ("b3_t", _I), ("a3_t", _I),
        ("n3_e1", _I), ("v3_e1", _I), ("b3_e1", _I), ("a3_e1", _I),
        ("w2_t", _PI), ("w2_e1", _PI),
        ("n2_t", _I), ("v2_t", _I), ("b2_t", _I), ("a2_t", _I),
        ("n2_e1", _I), ("v2_e1", _I), ("b2_e1", _I), ("a2_e1", _I),
        ("nbpt", _I), ("bpt", _PI), ("bin", _PI),
        ("ngsed", _I), ("gsed_t", _PI), ("ised_t", _PI),
        ("sur_t", _PI), ("nsur_t", _PI), ("bot_t", _PI),
        ("h1au1", _PD), ("h2au1", _PD), ("h1acell", _PD), ("cellarea", _PD),
        ("costhu1", _PD), ("sinthu1", _PD), ("cellx", _PD), ("celly", _PD),
        ("botz", _PD), ("gridz", _PD), ("cellz", _PD),
        ("szv", _I), ("szvS", _I), ("nvcm", _I),
        ("c2v", _PPI), ("v2c", _PPI), ("nvc", _PI), ("m2dv", _PI), ("e2v", _PPI),
        ("gridx", _PD), ("gridy", _PD), ("u1x", _PD), ("u1y", _PD),
        ("costhu2", _PD), ("sinthu2", _PD), ("wgst", _PI),
    ]


class QWeights(C.Structure):
    _fields_ = [
        ("B", _PPD), ("w", _D * 6), ("ncells", _I), ("cells", _PI),
        ("tmx", _D), ("tmn", _D), ("orf", _I), ("xr", _D), ("yr", _D),
        ("beta", _D), ("alpha", _D),
    ]


class Window(C.Structure):
    _fields_ = [
        ("ntr", _I), ("tr_wc", _PPD), ("dt", _D), ("dttr", _D), ("t", _D),
        ("nstep", _I),
        ("u1", _PD), ("u1flux3d", _PD), ("dzu1", _PD), ("u", _PD), ("v", _PD),
        ("w", _PD), ("Kz", _PD), ("eta", _PD), ("topz", _PD),
        ("u2", _PD), ("wtop", _PD), ("wbot", _PD),
    ]


class WinPriv(C.Structure):
    _fields_ = [
        ("trasc", _I), ("ultimate", _I), ("stab", C.c_int), ("sigma", _I),
        ("thin_merge", _I), ("rkstage", _I), ("means", _I),
        ("hmin", _D), ("trsf", _D),
        ("advect", _PI), ("diffuse", _PI), ("tbdy", _PI), ("ntbdy", _I),
        ("tdif_v", _PI), ("ntdif_v", _I), ("mintr", _PD), ("maxtr", _PD),
        ("oldeta", _PD), ("dz", _PD), ("Ds", _PD), ("Hn1", _PD), ("Hs", _PD),
        ("u1kh", _PD), ("u2kh", _PD), ("mdx", _PD), ("one", _PD),
        ("s1", _PI), ("vc", _I), ("vcs", _I), ("vc1", _I), ("vcs1", _I),
        ("i1", _PI), ("i2", _PI), ("i3", _PI), ("i4", _PI), ("i5", _PI),
    ] + [("w%d" % i, _PD) for i in range(1, 11)] + \
        [("d%d" % i, _PD) for i in range(1, 5)] + \
        [("v%d" % i, _PD) for i in range(1, 8)] + [("b1", _D)] + \
        [("flag", _I), ("fillf", _I), ("compatible", _I), ("lw", C.POINTER(QWeights)),
         ("twin", _PI)]


class Master(C.Structure):
    _fields_ = [
        ("mode2d", _I), ("crf", _I), ("win_data_empty_3d", C.c_void_p),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("window", C.POINTER(Geometry)), ("windat", C.POINTER(Window)),
        ("wincon", C.POINTER(WinPriv)),
        ("nmc", _I), ("npem", _I), ("nz", _I), ("ntr", _I),
        ("kbot", _PI), ("idx3", _PI), ("gidx", _PI), ("eidx", _PI), ("e2m", _PI),
        ("sedg", _PI), ("nv", _I), ("v2m", _PI),
    ]


def view(ptr, n, dtype):
    """numpy view of n elements of a C array (no copy)."""
    ct = {np.float64: C.c_double, np.int32: C.c_int}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,))


def view2(pptr, rows, cols):
    """numpy view of an int** allocated as one block of rows x cols."""
    return np.ctypeslib.as_array(C.cast(pptr[0], C.POINTER(C.c_int)), shape=(rows, cols))


class VLsq(C.Structure):
    """emstr_us_vlsq: the streamline velocity fit's weights of a window."""
    _fields_ = [("szc", C.c_int), ("nent", C.c_int),
                ("vo", C.POINTER(C.c_int)), ("vn", C.POINTER(C.c_int)),
                ("cell", C.POINTER(C.c_int)), ("edge", C.POINTER(C.c_int)),
                ("bx", C.POINTER(C.c_double)), ("by", C.POINTER(C.c_double))]
