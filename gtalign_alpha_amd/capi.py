"""ctypes mirror of include/gta/gtalign_amd.h (the C ABI of the stage)."""
import ctypes as C

GTA_TFM_NFIELDS = 12
GTA_ALN_NFIELDS = 15
GTA_ALN_NROWS = 5

# field indices of the per-pair alignment record (gproc.h:207-227)
(ALN_ORGSTRNO, ALN_STRNEWDST, ALN_BEGCOORDS, ALN_ENDCOORDS, ALN_ALNLENGTH, ALN_PSTVS,
 ALN_IDNTS, ALN_NGAPS, ALN_RMSD, ALN_SCOREQ, ALN_SCORER, ALN_D0Q, ALN_D0R,
 ALN_2SCOREQ, ALN_2SCORER) = range(15)

DEPTH_DEEP, DEPTH_HIGH, DEPTH_MEDIUM, DEPTH_SHALLOW = range(4)
REF_LOG, REF_COARSE, REF_ONE, REF_FULL, REF_FULLA = range(5)
PREC_FP32, PREC_MIXED = 0, 1

STATUS = {
    0: "GTA_OK", -1: "GTA_ERR_ARG", -2: "GTA_ERR_EMPTY", -3: "GTA_ERR_TOOMANY",
    -4: "GTA_ERR_SIZE", -5: "GTA_ERR_LENGTH", -6: "GTA_ERR_HIP", -7: "GTA_ERR_STATE",
}
STATUS_CODES = {v: k for k, v in STATUS.items()}


class ChunkView(C.Structure):
    _fields_ = [
        ("nstrs", C.c_int), ("nposs", C.c_int),
        ("len", C.POINTER(C.c_int)), ("dst", C.POINTER(C.c_uint)), ("type", C.POINTER(C.c_int)),
        ("x", C.POINTER(C.c_float)), ("y", C.POINTER(C.c_float)), ("z", C.POINTER(C.c_float)),
        ("ins_ch_ord", C.POINTER(C.c_int)), ("resnum", C.POINTER(C.c_int)),
        ("rsd", C.c_char_p), ("ss", C.c_char_p),
    ]


class Options(C.Structure):
    _fields_ = [
        ("prescore", C.c_float), ("scorethld", C.c_float), ("convergence", C.c_int),
        ("depth", C.c_int), ("refinement", C.c_int), ("symmetric", C.c_int),
        ("nodeletions", C.c_int), ("referenced", C.c_int), ("precision", C.c_int),
        ("batched", C.c_int),
        ("addsearchbyss", C.c_int),
        ("qgroup", C.c_int),
    ]


class Results(C.Structure):
    _fields_ = [
        ("nqrys", C.c_int), ("nrefs", C.c_int),
        ("alndata", C.POINTER(C.c_float)), ("tfms", C.POINTER(C.c_float)),
        ("alns", C.POINTER(C.c_char)), ("alnsize", C.c_size_t),
        ("seconds", C.c_double), ("dp_cells", C.c_double),
    ]


def default_options(**kw):
    """Reference defaults (CLOptions.cpp:35-100) with keyword overrides."""
    o = Options(prescore=0.4, scorethld=0.5, convergence=18, depth=DEPTH_MEDIUM,
                refinement=REF_ONE, symmetric=0, nodeletions=0, referenced=0,
                precision=PREC_FP32, batched=0, addsearchbyss=0, qgroup=0)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def alignment_section_size(nq, nqposs, nr, nrposs):
    """Bytes of the alignment section (CuMemoryBase::GetSizeOfDP2Alns)."""
    return (nrposs * nq + nr * (nqposs + nq)) * GTA_ALN_NROWS
