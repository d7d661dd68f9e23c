"""ctypes binding of the whole-scan C entry points (dwba_scan_* in
include/dwba_hip.h): the scan is planned, laid out in device memory,
launched and captured by the library's C++ engine; Python only passes the
host arrays.  It replaces the same per-pair loops as ScanBatch
(driver.py:907-1029, ionization.py:991-1002) with the FP64 step ScanBatch's
tuner picks by default (tail stage forked onto a second stream, shared
direct beta per energy, split-lane contraction, grouped Filon prep).

    scan = CScan(energies, L_max=30, oscillatory_method="full_split")
    scan.launch()                 # or: scan.capture(); scan.replay()
    I = scan.results()            # (n_pairs, 2, L_max+1): [direct, exchange]
    scan.close()

`energies` are EnergyInput-like objects (grid, V_core, U_i, U_f, bound_i,
bound_f, pairs of (cont_i, cont_f) or (cont_i, cont_f, bound_f)).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip


class CScan:
    def __init__(self, energies, L_max: int, use_oscillatory_quadrature=True, oscillatory_method="advanced",
                 CC_nodes=5, phase_increment=1.5708, min_grid_fraction=0.10, k_threshold=0.5):
        if oscillatory_method not in _hip.SCAN_METHODS:
            raise ValueError(f"Unknown oscillatory_method: {oscillatory_method}")
        self.lib = _hip.load_library()
        _hip.ensure_init()
        self.L_max = int(L_max)
        keep = []

        def arr(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.ndim != 1:
                raise ValueError("arrays must be one-dimensional")
            keep.append(a)
            return a.ctypes.data, len(a)

        E = (_hip.ScanEnergy * len(energies))()
        waves, wave_id, pairs, uf = [], {}, [], {}
        for e, en in enumerate(energies):
            g = en.grid
            r, n = arr(g.r)
            for name, a in (("w_simpson", g.w_simpson), ("w_trapz", g.w_trapz), ("V_core", en.V_core),
                            ("U_i", en.U_i), ("bound_i.u", en.bound_i.u_of_r)):
                if len(a) < n:
                    raise ValueError(f"{name} has {len(a)} pts but grid has {n}")
            U_f = arr(en.U_f)[0] if getattr(en, "U_f", None) is not None else None
            E[e] = _hip.ScanEnergy(r, arr(g.w_simpson)[0], arr(g.w_trapz)[0], arr(en.V_core)[0], arr(en.U_i)[0], U_f,
                                   arr(en.bound_i.u_of_r)[0], n, 0)
            for pr in en.pairs:
                bf = pr[2] if len(pr) > 2 else en.bound_f
                ids = []
                for w in pr[:2]:
                    key = (e, id(w))
                    if key not in wave_id:
                        p, m = arr(w.chi_of_r)
                        wave_id[key] = len(waves)
                        waves.append(_hip.ScanWave(p, float(w.k_au), float(getattr(w, "phase_shift", 0.0)),
                                                   float(getattr(w, "eta", 0.0)), float(getattr(w, "sigma_l", 0.0)),
                                                   int(w.l), int(getattr(w, "idx_match", -1)), m, 0))
                    ids.append(wave_id[key])
                key = (e, id(bf))
                if key not in uf:
                    uf[key] = arr(bf.u_of_r)
                p, m = uf[key]
                pairs.append(_hip.ScanPair(p, m, e, ids[0], ids[1], int(not hasattr(bf, "k_au")), 0))
        self.n_pairs = len(pairs)
        W = (_hip.ScanWave * len(waves))(*waves)
        P = (_hip.ScanPair * len(pairs))(*pairs)
        opt = _hip.ScanOptions(self.L_max, int(bool(use_oscillatory_quadrature)),
                               _hip.SCAN_METHODS[oscillatory_method], int(CC_nodes), float(phase_increment),
                               float(min_grid_fraction), float(k_threshold))
        h = ctypes.c_void_p()
        rc = self.lib.dwba_scan_plan(len(E), E, len(W), W, len(P), P, ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise (ValueError if rc == 1 else RuntimeError)(f"dwba_scan_plan: {_hip.last_error()}")
        self.handle = h
        self._capture_stream = None

    def _h(self):
        if not self.handle:
            raise RuntimeError("scan closed")
        return self.handle

    def launch(self, stream=None):
        _hip.check(self.lib.dwba_scan_launch(self._h(), _hip.stream_handle(stream)), "dwba_scan_launch")

    def capture(self, stream=None):
        """Captures the launch sequence into a HIP graph on `stream` (a
        created stream; a private one by default)."""
        import torch
        if stream is None:
            if self._capture_stream is None:
                self._capture_stream = torch.cuda.Stream()
            stream = self._capture_stream
        _hip.check(self.lib.dwba_scan_capture(self._h(), _hip.stream_handle(stream)), "dwba_scan_capture")

    def replay(self, stream=None):
        _hip.check(self.lib.dwba_scan_replay(self._h(), _hip.stream_handle(stream)), "dwba_scan_replay")

    def results(self, stream=None) -> np.ndarray:
        out = np.empty((self.n_pairs, 2, self.L_max + 1))
        _hip.check(self.lib.dwba_scan_read_results(self._h(), out.ctypes.data, _hip.stream_handle(stream)),
                   "dwba_scan_read_results")
        return out

    def info(self) -> dict:
        v = (ctypes.c_int64 * 8)()
        _hip.check(self.lib.dwba_scan_info(self._h(), v, 8), "dwba_scan_info")
        keys = ("pairs", "L_max", "device_bytes", "tail_pairs", "filon_groups", "shared_beta_energies",
                "coulomb_node_buffers", "moment_sets")
        return dict(zip(keys, (int(x) for x in v)))

    def close(self):
        if self.handle:
            self.lib.dwba_scan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
