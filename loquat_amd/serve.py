"""Front end of the C++ request loop (lq_server_*): submit runs, tick, collect
completions.  No torch needed; with fake=True nothing touches a GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .cases import Case


class Server:
    def __init__(self, np_cap: int, ncell_cap: int, max_runs: int, hbm_budget: int, steps_per_tick: int,
                 fake: bool = False, keep_results: bool = False):
        cfg = abi.LqServerConfig(int(np_cap), int(ncell_cap), int(max_runs), int(hbm_budget), int(steps_per_tick),
                                 1 if fake else 0, 1 if keep_results else 0)
        self.h = C.c_void_p()
        abi.check(abi.lib().lq_server_create(C.byref(cfg), C.byref(self.h)))
        self._buf = (C.c_int32 * 4096)()

    @staticmethod
    def footprint(case: Case) -> int:
        """Device bytes the run occupies once admitted."""
        p = case.derived()
        return abi.lib().lq_engine_bytes_per_run(p.np, p.ncell)

    def submit(self, case: Case, nsteps: int | None = None) -> int:
        ri = case.run_input(nsteps)
        rid = C.c_int32()
        abi.check(abi.lib().lq_server_submit(self.h, C.byref(ri), C.byref(rid)))
        return rid.value

    def tick(self) -> list[int]:
        n = C.c_int32()
        abi.check(abi.lib().lq_server_tick(self.h, self._buf, len(self._buf), C.byref(n)))
        return [self._buf[i] for i in range(min(n.value, len(self._buf)))]

    def info(self, rid: int) -> dict:
        st, done, ns, b = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        abi.check(abi.lib().lq_server_request_info(self.h, rid, C.byref(st), C.byref(done), C.byref(ns), C.byref(b)))
        return dict(state=abi.REQ_STATE[st.value], steps=done.value, nsteps=ns.value, bytes=b.value)

    def stats(self) -> dict:
        q, r, b, m = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        abi.check(abi.lib().lq_server_stats(self.h, C.byref(q), C.byref(r), C.byref(b), C.byref(m)))
        return dict(queued=q.value, running=r.value, bytes=b.value, npm=m.value)

    def result(self, rid: int, n: int) -> dict:
        out = {k: np.zeros((n, w), np.float32) for k, w in
               (("pospres", 4), ("velrhop", 4), ("idep", 4), ("str", 6), ("stn", 6))}
        posd = np.zeros((n, 3), np.float64)
        hp = abi.host_particles(out["pospres"], out["velrhop"], out["idep"], out["str"], out["stn"])
        abi.check(abi.lib().lq_server_result(self.h, rid, C.byref(hp), posd.ctypes.data))
        out["posd"] = posd
        return out

    def close(self) -> None:
        if self.h:
            abi.lib().lq_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
