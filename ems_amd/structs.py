"""ctypes mirror of include/emstr.h (geometry_t, window_t, win_priv_t).

Field order and types must match the C header exactly.  Field names and
meanings are those of the reference structures
(model/hd/include/sparse.h:246-654, 660-1004, 2393-2794).
"""
import ctypes as C

import numpy as np

_I = C.c_int
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PD = C.POINTER(C.c_double)


class OpenBdry(C.Structure):
    """open_bdrys_t (include/emstr.h)"""
    _fields_ = [
        ("ocodex", _I), ("ocodey", _I), ("nmap", _PI), ("no3_t", _I), ("no2_t", _I),
        ("obc_t", _PI), ("oi1_t", _PI), ("oi2_t", _PI), ("cyc_t", _PI), ("bot_t", _PI),
        ("obc_e1", _PI), ("oi1_e1", _PI), ("obc_e2", _PI), ("oi1_e2", _PI),
        ("t_transfer", C.POINTER(_PD)), ("bgz", _I), ("trm", _PI), ("upmeth", _I),
        ("bcond_tra", _PI), ("clampv", _PD), ("trpc", _PD), ("relax_zone_tra", _PI),
        ("options", _I),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("sgnum", _I), ("sgnumS", _I), ("sgsiz", _I), ("sgsizS", _I),
        ("nobc", _I), ("open", C.POINTER(C.POINTER(OpenBdry))),
        ("nce1", _I), ("nce2", _I), ("nz", _I), ("wn", _I),
        ("layers", _PD),
        ("s2i", _PI), ("s2j", _PI), ("s2k", _PI),
        ("w3_t", _PI), ("w3_e1", _PI), ("w3_e2", _PI),
        ("n3_t", _I), ("n3_e1", _I), ("n3_e2", _I),
        ("v3_t", _I), ("v3_e1", _I), ("v3_e2", _I),
        ("b3_t", _I), ("b3_e1", _I), ("b3_e2", _I),
        ("a3_t", _I), ("a3_e1", _I), ("a3_e2", _I),
        ("w2_t", _PI), ("w2_e1", _PI), ("w2_e2", _PI),
        ("n2_t", _I), ("n2_e1", _I), ("n2_e2", _I),
        ("v2_t", _I), ("v2_e1", _I), ("v2_e2", _I),
        ("b2_t", _I), ("b2_e1", _I), ("b2_e2", _I),
        ("a2_t", _I), ("a2_e1", _I), ("a2_e2", _I),
        ("nbpt", _I), ("bpt", _PI), ("bin", _PI),
        ("ngsed", _I), ("gsed_t", _PI), ("ised_t", _PI),
        ("sur_t", _PI), ("nsur_t", _PI), ("bot_t", _PI), ("m2d", _PI),
        ("xp1", _PI), ("xm1", _PI), ("yp1", _PI), ("ym1", _PI),
        ("zp1", _PI), ("zm1", _PI),
        ("h1acell", _PD), ("h2acell", _PD), ("h1au1", _PD), ("h2au2", _PD),
        ("h2au1", _PD), ("h1au2", _PD), ("cellarea", _PD),
        ("botz", _PD), ("gridz", _PD), ("cellz", _PD),
        ("sur_e1", _PI), ("sur_e2", _PI),
        ("botzu1", _PD), ("botzu2", _PD), ("bot_e1", _PI), ("bot_e2", _PI),
    ]


class TracerInfo(C.Structure):
    """tracer_info_t of include/emstr.h (the fields auxiliary_routines
    reads)"""
    _fields_ = [
        ("units", C.c_char * 64), ("flag", _I), ("relax_rate", _D),
        ("relax_dum", _D), ("tctype", _I),
    ]


class Window(C.Structure):
    _fields_ = [
        ("ntr", _I), ("tr_wc", C.POINTER(_PD)), ("dt", _D), ("dttr", _D),
        ("t", _D), ("nstep", _I),
        ("u1", _PD), ("u2", _PD), ("w", _PD),
        ("u1flux3d", _PD), ("u2flux3d", _PD),
        ("dzu1", _PD), ("dzu2", _PD), ("Kz", _PD),
        ("eta", _PD), ("topz", _PD), ("wbot", _PD),
        ("u1flux", _PD), ("u2flux", _PD), ("u1vm", _PD), ("u2vm", _PD),
        ("etab", _PD), ("wtop", _PD), ("detadt", _PD),
        ("tno", _I), ("sno", _I), ("rtemp", _PD), ("rsalt", _PD),
        ("temp_tc", _PD), ("salt_tc", _PD), ("agetr", _PD),
    ]


class WinPriv(C.Structure):
    _fields_ = [
        ("trasc", _I), ("ultimate", _I), ("stab", _I), ("sigma", _I),
        ("thin_merge", _I), ("hmin", _D),
        ("advect", _PI), ("diffuse", _PI), ("tbdy", _PI), ("ntbdy", _I),
        ("tdif_v", _PI), ("ntdif_v", _I), ("mintr", _PD), ("maxtr", _PD),
        ("oldeta", _PD), ("dz", _PD), ("Ds", _PD), ("Hn1", _PD), ("Hs", _PD),
        ("u1kh", _PD), ("u2kh", _PD), ("mdx", _PD), ("mdy", _PD), ("one", _PD),
        ("s1", _PI), ("vc", _I), ("vcs", _I), ("vc1", _I), ("vcs1", _I),
        ("i1", _PI), ("i2", _PI), ("i3", _PI), ("i4", _PI), ("i5", _PI),
        ("w1", _PD), ("w2", _PD), ("w3", _PD), ("w4", _PD), ("w5", _PD),
        ("w6", _PD), ("w7", _PD), ("w8", _PD), ("w9", _PD), ("w10", _PD),
        ("d1", _PD), ("d2", _PD), ("d3", _PD), ("d4", _PD),
        ("v1", _PD), ("v2", _PD), ("v3", _PD), ("v4", _PD), ("v5", _PD),
        ("v6", _PD), ("v7", _PD),
        ("b1", _D), ("twin", _PI),
        ("cbot_e1", _PI), ("cbot_e2", _PI), ("ncbot_e1", _I), ("ncbot_e2", _I),
        ("nonlinear", _I),
        ("ntdec", _I), ("tdec", _PI), ("dectr", _PD),
        ("trinfo_3d", C.POINTER(TracerInfo)), ("agemsk", C.POINTER(C.c_short)),
        ("s2", _PI), ("vc2", _I), ("vcs2", _I), ("vca2", _I), ("vci2", _I),
        ("c1", C.POINTER(C.c_char)),
    ]


class Master(C.Structure):
    _fields_ = [
        ("mode2d", _I), ("crf", _I), ("win_data_empty_3d", C.c_void_p),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("window", C.POINTER(Geometry)), ("windat", C.POINTER(Window)),
        ("wincon", C.POINTER(WinPriv)),
        ("nx", _I), ("ny", _I), ("nz", _I), ("ntr", _I), ("kbot", _PI),
    ]


# Scheme / stability flags (model/hd/include/hd_params.h:104, 236-242, 331-350)
ORDER1 = 0x000002
ORDER2 = 0x000004
ORDER4 = 0x000010
QUICKEST = 0x000020
QUICKEST_CO = 0x000040
VANLEER = 0x000080
ORDER2_UW = 0x002000
FFSL = 0x080000
# open boundaries (include/emstr.h)
L_EDGE, R_EDGE, B_EDGE, F_EDGE = 0x80, 0x100, 0x400, 0x800
FILEIN, CUSTOM, NOGRAD, LINEXT, POLEXT = 0x1, 0x2, 0x4, 0x8, 0x10
CYCLIC, NOTHIN, CLAMPD, STATIS, UPSTRM = 0x400, 0x800, 0x4000, 0x8000, 0x10000
TRFLUX, TRCONC, TRCONF = 0x1000000, 0x2000000, 0x4000000
OP_OWRITE, OP_OBCCM = 0x20000, 0x40000
CENTER, FACE, ADAPTIVE = 1, 2, 8
SUB_STEP = 2
TMASS = 16
TRACERS = 256
RS_RESTART = 0x0004
DE_TR2 = 0x0400
DE_TR3 = 0x0800
RLX_ADPT = 0x0008
RLX_GRD = 0x0400
RLX_REG = 0x0800
RLX_OBC = 0x1000
NONE = 0x80000000 - (1 << 32)  # as a signed C int


def view(ptr, n, dtype):
    """numpy view (no copy) of n elements behind a ctypes pointer."""
    if not ptr:
        raise ValueError("null pointer")
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)
