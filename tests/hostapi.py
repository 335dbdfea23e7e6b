"""ctypes mirror of include/graphblox_hip.h (the reference host API layouts)."""
import ctypes as C
import os

import numpy as np

from conftest import ROOT

u32 = C.c_uint32


class MT(C.Structure):
    _fields_ = [("mt", C.c_ulong * 624), ("mti", C.c_int)]


class Arguments(C.Structure):
    _fields_ = [("wflag", C.c_int), ("xflag", C.c_int), ("sflag", C.c_int), ("Sflag", C.c_int),
                ("dflag", C.c_int), ("binSize", u32), ("verbosity", u32), ("iterations", u32),
                ("trials", u32), ("epsilon", C.c_double), ("source", u32), ("algorithm", u32),
                ("datastructure", u32), ("pushpull", u32), ("sort", u32), ("lmode", u32),
                ("lmode_l2", u32), ("lmode_l3", u32), ("mmode", u32), ("symmetric", u32),
                ("weighted", u32), ("delta", u32), ("pre_numThreads", C.c_int),
                ("algo_numThreads", C.c_int), ("ker_numThreads", C.c_int), ("fnameb", C.c_char_p),
                ("fnamel", C.c_char_p), ("fnameb_format", u32), ("convert_format", u32),
                ("mt19937var", MT), ("cache_size", u32), ("kernel_name", C.c_char_p), ("cu_id", C.c_int),
                ("device_index", C.c_int), ("xclbin_path", C.c_char_p), ("overlay_path", C.c_char_p),
                ("ctrl_mode", C.c_int), ("endian_read", C.c_int), ("endian_write", C.c_int),
                ("flush_cache", C.c_int), ("bankGroupIndex", C.c_int)]


class EdgeList(C.Structure):
    _fields_ = [("num_edges", u32), ("num_vertices", u32), ("avg_degree", u32), ("max_weight", C.c_float),
                ("edges_array_weight", C.POINTER(C.c_float)), ("edges_array_src", C.POINTER(u32)),
                ("edges_array_dest", C.POINTER(u32)), ("mask_array", C.POINTER(u32)),
                ("label_array", C.POINTER(u32)), ("inverse_label_array", C.POINTER(u32))]


class Vertex(C.Structure):
    _fields_ = [("num_vertices", u32), ("out_degree", C.POINTER(u32)), ("in_degree", C.POINTER(u32)),
                ("edges_idx", C.POINTER(u32))]


def graph_csr_type(weighted: bool, directed: bool = True):
    """struct GraphCSR in the (DIRECTED, WEIGHTED) layout (graphCSR.h:14-31)."""
    fields = [("num_edges", u32), ("num_vertices", u32), ("avg_degree", u32)]
    if weighted:
        fields.append(("max_weight", C.c_float))
    fields += [("vertices", C.POINTER(Vertex)), ("sorted_edges_array", C.POINTER(EdgeList))]
    if directed:
        fields += [("inverse_vertices", C.POINTER(Vertex)), ("inverse_sorted_edges_array", C.POINTER(EdgeList))]
    name = "GraphCSR" + ("W" if weighted else "") + ("" if directed else "D0")
    return type(name, (C.Structure,), {"_fields_": fields})


class PageRankStats(C.Structure):
    _fields_ = [("damp", C.c_float), ("base_pr", C.c_float), ("iterations", u32), ("num_vertices", u32),
                ("activeVertices", u32), ("realRanks", C.POINTER(u32)), ("pageRanks", C.POINTER(C.c_float)),
                ("time_total", C.c_double), ("error_total", C.c_double)]


class SPMVStats(C.Structure):
    _fields_ = [("iterations", u32), ("arguments", C.POINTER(Arguments)), ("num_vertices", u32),
                ("vector_output", C.POINTER(C.c_float)), ("vector_input", C.POINTER(C.c_float)),
                ("time_total", C.c_double)]


def library_name(weighted: bool = False, directed: bool = True) -> str:
    return "libgraphblox_hip" + ("" if directed else "_d0") + ("_w" if weighted else "") + ".so"


def load(weighted: bool = False, directed: bool = True):
    L = C.CDLL(os.path.join(ROOT, "graphblox_amd", library_name(weighted, directed)))
    G = graph_csr_type(weighted, directed)
    pA, pG = C.POINTER(Arguments), C.POINTER(G)
    L.argumentsNew.restype = pA
    L.argumentsFree.argtypes = [pA]
    L.readEdgeListsbin.restype = C.POINTER(EdgeList)
    L.readEdgeListsbin.argtypes = [C.c_char_p, C.c_uint8, u32, u32]
    L.freeEdgeList.argtypes = [C.POINTER(EdgeList)]
    L.graphCSRPreProcessingStep.restype = pG
    L.graphCSRPreProcessingStep.argtypes = [pA]
    L.graphCSRFree.argtypes = [pG]
    L.pageRankGraphCSR.restype = C.POINTER(PageRankStats)
    L.pageRankGraphCSR.argtypes = [pA, C.c_void_p]
    L.pageRankPullGraphCSR.restype = C.POINTER(PageRankStats)
    L.pageRankPullGraphCSR.argtypes = [pA, C.c_void_p]
    L.freePageRankStats.argtypes = [C.POINTER(PageRankStats)]
    L.SPMVGraphCSR.restype = C.POINTER(SPMVStats)
    L.SPMVGraphCSR.argtypes = [pA, C.c_void_p]
    L.SPMVPullGraphCSR.restype = C.POINTER(SPMVStats)
    L.SPMVPullGraphCSR.argtypes = [pA, C.c_void_p]
    L.freeSPMVStats.argtypes = [C.POINTER(SPMVStats)]
    L.graphbloxHipShutdown.argtypes = []
    L.graphbloxHipForget.argtypes = [C.c_void_p]
    L.runSPMVAlgorithm.restype = C.POINTER(SPMVStats)
    L.runSPMVAlgorithm.argtypes = [pA, C.c_void_p]
    L.GraphCSR = G
    return L


class HostGraph:
    """A GraphCSR assembled in Python (numpy arrays, the reference's struct
    layout), as code other than this library's graphCSRPreProcessingStep
    would hand one over.  `keep` holds every array the struct points to."""

    def __init__(self, L, N, E, fwd, inv=None, weighted=False):
        self.L = L
        self.keep = []
        self.g = L.GraphCSR()
        self.set(N, E, fwd, inv, weighted)

    def _vertex(self, N, a):
        v = Vertex()
        v.num_vertices = N
        for nm in ("out_degree", "in_degree", "edges_idx"):
            arr = np.ascontiguousarray(a.get(nm, np.zeros(N, np.uint32)), np.uint32)
            self.keep.append(arr)
            setattr(v, nm, arr.ctypes.data_as(C.POINTER(u32)))
        self.keep.append(v)
        return C.pointer(v)

    def _edges(self, N, E, a):
        e = EdgeList()
        e.num_edges, e.num_vertices, e.avg_degree, e.max_weight = E, N, E // max(N, 1), 1.0
        for nm, key, dt in (("edges_array_src", "src", np.uint32), ("edges_array_dest", "dst", np.uint32),
                            ("edges_array_weight", "w", np.float32)):
            arr = np.ascontiguousarray(a.get(key, np.ones(E, dt)), dt)
            self.keep.append(arr)
            ct = C.c_float if dt == np.float32 else u32
            setattr(e, nm, arr.ctypes.data_as(C.POINTER(ct)))
        self.keep.append(e)
        return C.pointer(e)

    def set(self, N, E, fwd, inv=None, weighted=False):
        """Point the same struct (same address) at new arrays."""
        g = self.g
        g.num_edges, g.num_vertices, g.avg_degree = E, N, E // max(N, 1)
        if weighted:
            g.max_weight = 1.0
        g.vertices = self._vertex(N, fwd)
        g.sorted_edges_array = self._edges(N, E, fwd)
        if inv is not None:
            g.inverse_vertices = self._vertex(N, inv)
            g.inverse_sorted_edges_array = self._edges(N, E, inv)

    @property
    def ptr(self):
        return C.addressof(self.g)
