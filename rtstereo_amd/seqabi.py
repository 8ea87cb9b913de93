"""ctypes binding of a library that exports the reference's seq/ stage
functions under their C++ (Itanium-mangled) names: librtstereo_seq.so, or a
build of the reference's own seq/ sources.  The wrappers take numpy arrays in
the reference's host layouts and fill outputs in place, exactly as a C++
caller of those functions would.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

F, I, P = C.c_float, C.c_int, C.c_void_p

# (mangled name, argtypes, restype)
SYMBOLS = {
    "pixelBasedL2R": ("_Z13pixelBasedL2RPhS_Pfiii", [P, P, P, I, I, I], None),
    "pixelBasedR2L": ("_Z13pixelBasedR2LPhS_Pfiii", [P, P, P, I, I, I], None),
    "pixelBasedCostL2R_Float": ("_Z23pixelBasedCostL2R_FloatPfS_S_iii", [P, P, P, I, I, I], None),
    "pixelBasedCostR2L_Float": ("_Z23pixelBasedCostR2L_FloatPfS_S_iii", [P, P, P, I, I, I], None),
    "pixelBasedCostL2R_Color": ("_Z23pixelBasedCostL2R_ColorPhS_Pfiiii", [P, P, P, I, I, I, I], None),
    "pixelBasedCostL2R_Color_trunc": ("_Z23pixelBasedCostL2R_ColorPhS_Pfiiifi", [P, P, P, I, I, I, F, I], None),
    "pixelBasedR2L_Color": ("_Z19pixelBasedR2L_ColorPhS_Pfiiii", [P, P, P, I, I, I, I], None),
    "pixelBasedR2L_Color_trunc": ("_Z19pixelBasedR2L_ColorPhS_Pfiiifi", [P, P, P, I, I, I, F, I], None),
    "pixelBasedCostR2L_Color": ("_Z23pixelBasedCostR2L_ColorPhS_Pfiiii", [P, P, P, I, I, I, I], None),
    "pixelBasedCostR2L_Color_trunc": ("_Z23pixelBasedCostR2L_ColorPhS_Pfiiifi", [P, P, P, I, I, I, F, I], None),
    "calcCensusTrans": ("_Z15calcCensusTransPhPbiiii", [P, P, I, I, I, I], None),
    "calcHammingDist": ("_Z15calcHammingDistPbS_Pfiiiii", [P, P, P, I, I, I, I, I], None),
    "calculateRankFilt": ("_Z17calculateRankFiltPhS_iii", [P, P, I, I, I], None),
    "convolveImage": ("_Z13convolveImagePhPfS0_iiii", [P, P, P, I, I, I, I], None),
    "calcNCCDispL2R": ("_Z14calcNCCDispL2RPhS_Pfiiiii", [P, P, P, I, I, I, I, I], None),
    "combineCost": ("_Z11combineCostPfS_ffiii", [P, P, F, F, I, I, I], None),
    "boxAggregate": ("_Z12boxAggregatePfiiii", [P, I, I, I, I], None),
    "calcProxWeight": ("_Z14calcProxWeightPffi", [P, F, I], None),
    "calcEuclideanDist3P": ("_Z19calcEuclideanDist3Piiiiii", [I, I, I, I, I, I], F),
    "calcEuclideanDist3P_float": ("_Z25calcEuclideanDist3P_floatffffff", [F, F, F, F, F, F], F),
    "computeWeights": ("_Z14computeWeightsPfS_Phfiii", [P, P, P, F, I, I, I], None),
    "calcAWCostL2R": ("_Z13calcAWCostL2RPfS_S_S_iiii", [P, P, P, P, I, I, I, I], None),
    "calcAWCostR2L": ("_Z13calcAWCostR2LPfS_S_iiii", [P, P, P, I, I, I, I], None),
    "calc_x_range": ("_Z12calc_x_rangePhiiiiPiS0_", [P, I, I, I, I, P, P], None),
    "calc_y_range": ("_Z12calc_y_rangePhiiiiPiS0_", [P, I, I, I, I, P, P], None),
    "findCross": ("_Z9findCrossPhPiS0_ii", [P, P, P, I, I], None),
    "aggregate_cost_horizontal": ("_Z25aggregate_cost_horizontalPfiii", [P, I, I, I], None),
    "cross_stereo_aggregation": ("_Z24cross_stereo_aggregationPfS_PiS0_S0_S0_iii", [P, P, P, P, P, P, I, I, I], None),
    "disp_refinement": ("_Z15disp_refinementPhPiS0_iii", [P, P, P, I, I, I], None),
    "WTA": ("_Z3WTAPfPhiii", [P, P, I, I, I], None),
    "WTA_NCC": ("_Z7WTA_NCCPfPhiii", [P, P, I, I, I], None),
    "cross_check": ("_Z11cross_checkPhS_S_ii", [P, P, P, I, I], None),
}


class SeqLib:
    """Callable stage functions of one seq-ABI library."""

    def __init__(self, path: str):
        self.path = path
        self.lib = C.CDLL(path)
        self.fn = {}
        for name, (sym, args, res) in SYMBOLS.items():
            try:
                f = getattr(self.lib, sym)
            except AttributeError:
                continue
            f.argtypes = args
            f.restype = res
            self.fn[name] = f

    def has(self, name: str) -> bool:
        return name in self.fn

    @staticmethod
    def _p(a: np.ndarray):
        assert a.flags.c_contiguous
        return a.ctypes.data

    # Each wrapper allocates the outputs the way seq/Main.cpp does (malloc'd
    # host buffers) and returns them.
    def ad(self, name, ref, tar, D, *extra):
        H, W = ref.shape[:2]
        cost = np.zeros((H, W, D), np.float32)
        self.fn[name](self._p(ref), self._p(tar), self._p(cost), D, H, W, *extra)
        return cost

    def census(self, img, hr, wr):
        H, W = img.shape
        out = np.zeros((H, W, (2 * hr + 1) * (2 * wr + 1)), np.bool_)
        self.fn["calcCensusTrans"](self._p(img), self._p(out), hr, wr, H, W)
        return out

    def hamming(self, cref, ctar, hr, wr, D):
        H, W, _ = cref.shape
        cost = np.zeros((H, W, D), np.float32)
        self.fn["calcHammingDist"](self._p(cref), self._p(ctar), self._p(cost), hr, wr, H, W, D)
        return cost

    def rank(self, img, m):
        out = np.zeros_like(img)
        H, W = img.shape
        self.fn["calculateRankFilt"](self._p(img), self._p(out), m, H, W)
        return out

    def convolve(self, img, mask, hr, wr):
        H, W = img.shape
        out = np.zeros((H, W), np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        self.fn["convolveImage"](self._p(img), self._p(out), self._p(m), H, W, hr, wr)
        return out

    def ncc(self, ref, tar, D, hr, wr, init=None):
        H, W = ref.shape
        cost = np.zeros((H, W, D), np.float32) if init is None else np.ascontiguousarray(init, np.float32).copy()
        self.fn["calcNCCDispL2R"](self._p(ref), self._p(tar), self._p(cost), D, H, W, hr, wr)
        return cost

    def combine(self, c1, c2, l1, l2):
        H, W, D = c1.shape
        out = np.ascontiguousarray(c1, np.float32).copy()
        c2 = np.ascontiguousarray(c2, np.float32)
        self.fn["combineCost"](self._p(out), self._p(c2), l1, l2, H, W, D)
        return out

    def box(self, cost, r):
        H, W, D = cost.shape
        out = np.ascontiguousarray(cost, np.float32).copy()
        self.fn["boxAggregate"](self._p(out), H, W, D, r)
        return out

    def prox(self, gp, r):
        out = np.zeros(((2 * r + 1) ** 2,), np.float32)
        self.fn["calcProxWeight"](self._p(out), gp, r)
        return out

    def weights(self, prox, lab, gc, r):
        H, W, _ = lab.shape
        out = np.zeros((H, W, (2 * r + 1) ** 2), np.float32)
        prox = np.ascontiguousarray(prox, np.float32)
        self.fn["computeWeights"](self._p(out), self._p(prox), self._p(lab), gc, H, W, r)
        return out

    def aw_l2r(self, wref, wtar, cost, r):
        H, W, D = cost.shape
        out = np.zeros_like(cost)
        self.fn["calcAWCostL2R"](self._p(wref), self._p(wtar), self._p(np.ascontiguousarray(cost)), self._p(out),
                                 H, W, D, r)
        return out

    def find_cross(self, rgb):
        H, W, _ = rgb.shape
        horz = np.zeros((H, W, 2), np.int32)
        vert = np.zeros((H, W, 2), np.int32)
        self.fn["findCross"](self._p(rgb), self._p(horz), self._p(vert), H, W)
        return horz, vert

    def hprefix(self, cost):
        H, W, D = cost.shape
        out = np.ascontiguousarray(cost, np.float32).copy()
        self.fn["aggregate_cost_horizontal"](self._p(out), H, W, D)
        return out

    def cross_v(self, P_, href, htar, vref, vtar):
        H, W, D = P_.shape
        out = np.zeros_like(P_)
        self.fn["cross_stereo_aggregation"](self._p(np.ascontiguousarray(P_)), self._p(out), self._p(href),
                                            self._p(htar), self._p(vref), self._p(vtar), H, W, D)
        return out

    def refine(self, disp, horz, vert, D):
        H, W = disp.shape
        out = np.ascontiguousarray(disp).copy()
        self.fn["disp_refinement"](self._p(out), self._p(horz), self._p(vert), H, W, D)
        return out

    def wta(self, cost, ncc=False):
        H, W, D = cost.shape
        out = np.zeros((H, W), np.uint8)
        self.fn["WTA_NCC" if ncc else "WTA"](self._p(np.ascontiguousarray(cost)), self._p(out), H, W, D)
        return out

    def cross_check(self, dl, dr):
        H, W = dl.shape
        out = np.zeros((H, W), np.uint8)
        self.fn["cross_check"](self._p(dl), self._p(dr), self._p(out), H, W)
        return out
