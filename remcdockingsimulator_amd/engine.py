"""Engine: one device context of the C ABI with its systems and replicas.

This is the host-side object the Replica/Molecule facade and the tests drive.
Every compute call goes to libremc.so; nothing is computed here.
"""
import ctypes

import numpy as np

from . import _native as nat
from .forcefield import ForceField


class Engine:
    def __init__(self, device=0):
        self.lib = nat.load()
        h = ctypes.c_void_p()
        nat.check(self.lib, self.lib.remc_ctx_create(device, ctypes.byref(h)), "remc_ctx_create")
        self.ctx = h
        self.systems = []      # SystemArrays per system
        self.replica_sys = []  # system id per replica
        self._keep = []

    # ----------------------------------------------------------- setup ----
    def close(self):
        if self.ctx:
            self.lib.remc_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        nat.check(self.lib, rc, what)

    def set_forcefield(self, ff: ForceField):
        s = nat.remc_forcefield()
        flat = np.ascontiguousarray(ff.table, dtype=np.float32).reshape(-1)
        for i in range(flat.size):
            s.table[i] = flat[i]
        s.lambda_, s.e0, s.xi = ff.lambda_, ff.e0, ff.xi
        s.lj_cf, s.dh_cf, s.kbt_cf = ff.lj_cf, ff.dh_cf, ff.kbt_cf
        s.crowder_sigma, s.crowder_cutoff = ff.crowder_sigma, ff.crowder_cutoff
        self._chk(self.lib.remc_set_forcefield(self.ctx, ctypes.byref(s)), "remc_set_forcefield")

    def set_seed(self, seed):
        self._chk(self.lib.remc_set_seed(self.ctx, int(seed) & (2**64 - 1)), "remc_set_seed")

    def add_system(self, sa):
        d = nat.remc_system_desc()
        arrs = dict(
            body_start=np.ascontiguousarray(sa.body_start, dtype=np.int32),
            type=np.ascontiguousarray(sa.type, dtype=np.int32),
            charge=np.ascontiguousarray(sa.charge, dtype=np.float32),
            sigma=np.ascontiguousarray(sa.sigma, dtype=np.float32),
            crowder=np.ascontiguousarray(sa.crowder, dtype=np.uint8),
            x=np.ascontiguousarray(sa.pos[:, 0], dtype=np.float32),
            y=np.ascontiguousarray(sa.pos[:, 1], dtype=np.float32),
            z=np.ascontiguousarray(sa.pos[:, 2], dtype=np.float32),
            rx=np.ascontiguousarray(sa.rel[:, 0], dtype=np.float32),
            ry=np.ascontiguousarray(sa.rel[:, 1], dtype=np.float32),
            rz=np.ascontiguousarray(sa.rel[:, 2], dtype=np.float32),
            cx=np.ascontiguousarray(sa.center[:, 0], dtype=np.float32),
            cy=np.ascontiguousarray(sa.center[:, 1], dtype=np.float32),
            cz=np.ascontiguousarray(sa.center[:, 2], dtype=np.float32),
            quat=np.ascontiguousarray(sa.quat, dtype=np.float64).reshape(-1),
        )
        d.n_beads = sa.n_beads
        d.n_bodies = sa.n_bodies
        d.body_start = nat.iptr(arrs["body_start"])
        d.type = nat.iptr(arrs["type"])
        d.charge = nat.fptr(arrs["charge"])
        d.sigma = nat.fptr(arrs["sigma"])
        d.crowder = arrs["crowder"].ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        d.box = sa.box
        for k in ("x", "y", "z", "rx", "ry", "rz", "cx", "cy", "cz"):
            setattr(d, k, nat.fptr(arrs[k]))
        d.quat = nat.dptr(arrs["quat"])
        sid = ctypes.c_int()
        self._chk(self.lib.remc_system_add(self.ctx, ctypes.byref(d), ctypes.byref(sid)), "remc_system_add")
        self.systems.append(sa)
        return sid.value

    def add_replicas(self, sys_id, temperatures, translate_steps, rotate_steps):
        t = np.ascontiguousarray(temperatures, dtype=np.float64)
        ts = np.ascontiguousarray(np.broadcast_to(translate_steps, t.shape), dtype=np.float32)
        rs = np.ascontiguousarray(np.broadcast_to(rotate_steps, t.shape), dtype=np.float64)
        first = ctypes.c_int()
        self._chk(self.lib.remc_replicas_add(self.ctx, sys_id, len(t), nat.dptr(t), nat.fptr(ts), nat.dptr(rs),
                                             ctypes.byref(first)), "remc_replicas_add")
        self.replica_sys.extend([sys_id] * len(t))
        return first.value

    @property
    def n_replicas(self):
        return len(self.replica_sys)

    # --------------------------------------------------------- compute ----
    def energy_plain(self, replica, block=0):
        out = ctypes.c_double()
        self._chk(self.lib.remc_energy_plain(self.ctx, replica, block, ctypes.byref(out)), "remc_energy_plain")
        return out.value

    def energy_all(self):
        out = np.zeros(self.n_replicas, dtype=np.float64)
        self._chk(self.lib.remc_energy_all(self.ctx, nat.dptr(out)), "remc_energy_all")
        return out

    def energy_all_noalloc(self):
        """remc_energy_all_noalloc: the same energies from the library's static
        scratch (the context must be set up by an earlier call)."""
        out = np.zeros(self.n_replicas, dtype=np.float64)
        self._chk(self.lib.remc_energy_all_noalloc(self.ctx, nat.dptr(out)), "remc_energy_all_noalloc")
        return out

    def energy_noncrowder_all(self):
        out = np.zeros(self.n_replicas, dtype=np.float64)
        self._chk(self.lib.remc_energy_noncrowder_all(self.ctx, nat.dptr(out)), "remc_energy_noncrowder_all")
        return out

    def refresh_energies(self):
        self._chk(self.lib.remc_refresh_energies(self.ctx), "remc_refresh_energies")

    def mc_steps(self, nsteps):
        self._chk(self.lib.remc_mc_steps(self.ctx, int(nsteps)), "remc_mc_steps")

    def mc_steps_fb(self, nsteps, block=0):
        """Force-biased (smart) MC steps on every replica (remc_mc_steps_fb)."""
        self._chk(self.lib.remc_mc_steps_fb(self.ctx, int(nsteps), int(block)), "remc_mc_steps_fb")

    def mc_steps_plain(self, replica, nsteps, block=0):
        self._chk(self.lib.remc_mc_steps_plain(self.ctx, replica, int(nsteps), block), "remc_mc_steps_plain")

    def exchange(self):
        n = ctypes.c_int()
        self._chk(self.lib.remc_exchange(self.ctx, ctypes.byref(n)), "remc_exchange")
        return n.value

    def run(self, rounds, steps_per_round):
        self._chk(self.lib.remc_run(self.ctx, int(rounds), int(steps_per_round)), "remc_run")

    def set_run_refresh(self, enable=True):
        """Energy refresh inside run()'s captured rounds, before each exchange."""
        self._chk(self.lib.remc_set_run_refresh(self.ctx, 1 if enable else 0), "remc_set_run_refresh")

    def synchronize(self):
        self._chk(self.lib.remc_synchronize(self.ctx), "remc_synchronize")

    def set_recording(self, max_steps):
        self._chk(self.lib.remc_set_recording(self.ctx, int(max_steps)), "remc_set_recording")
        self._rec_max = int(max_steps)

    def records(self, replica, n=None):
        n = self._rec_max if n is None else n
        out = np.zeros(n, dtype=nat.RECORD_DTYPE)
        ptr = out.ctypes.data_as(ctypes.POINTER(nat.remc_step_record))
        self._chk(self.lib.remc_get_records(self.ctx, replica, ptr, n), "remc_get_records")
        return out

    def set_mc_block(self, block):
        self._chk(self.lib.remc_set_mc_block(self.ctx, int(block)), "remc_set_mc_block")

    def mc_block(self):
        b = ctypes.c_int()
        self._chk(self.lib.remc_get_mc_block(self.ctx, ctypes.byref(b)), "remc_get_mc_block")
        return b.value

    def fb_block(self):
        b = ctypes.c_int()
        self._chk(self.lib.remc_get_fb_block(self.ctx, ctypes.byref(b)), "remc_get_fb_block")
        return b.value

    def set_mc_coop(self, coop):
        self._chk(self.lib.remc_set_mc_coop(self.ctx, int(coop)), "remc_set_mc_coop")

    def mc_coop(self):
        b = ctypes.c_int()
        self._chk(self.lib.remc_get_mc_coop(self.ctx, ctypes.byref(b)), "remc_get_mc_coop")
        return b.value

    def set_mc_tuning(self, lds_cache=True, split_min_pairs=0):
        self._chk(self.lib.remc_set_mc_tuning(self.ctx, 1 if lds_cache else 0, int(split_min_pairs)),
                  "remc_set_mc_tuning")

    def set_mc_window(self, enable=True):
        """Speculative windows of small moves in the MC kernel (default on)."""
        self._chk(self.lib.remc_set_mc_window(self.ctx, 1 if enable else 0), "remc_set_mc_window")

    def adjoint(self, replica):
        m = self.systems[self.replica_sys[replica]].n_bodies
        ft = np.zeros((m, 6), dtype=np.float64)
        dp = np.zeros(3, dtype=np.float64)
        self._chk(self.lib.remc_adjoint(self.ctx, replica, nat.dptr(ft), nat.dptr(dp)), "remc_adjoint")
        return ft, dp

    def adjoint_all(self):
        """Forces/torques (list of (m,6) per replica) and dE/d(lambda,e0,xi)
        (R,3) of every replica from one batched launch."""
        sizes = [self.systems[s].n_bodies for s in self.replica_sys]
        ft = np.zeros((sum(sizes), 6), dtype=np.float64)
        dp = np.zeros((self.n_replicas, 3), dtype=np.float64)
        self._chk(self.lib.remc_adjoint_all(self.ctx, nat.dptr(ft), nat.dptr(dp)), "remc_adjoint_all")
        out, o = [], 0
        for m in sizes:
            out.append(ft[o:o + m])
            o += m
        return out, dp

    # ----------------------------------------------------------- state ----
    def energies(self):
        out = np.zeros(self.n_replicas, dtype=np.float64)
        self._chk(self.lib.remc_get_energies(self.ctx, nat.dptr(out)), "remc_get_energies")
        return out

    def temperatures(self):
        out = np.zeros(self.n_replicas, dtype=np.float64)
        self._chk(self.lib.remc_get_temperatures(self.ctx, nat.dptr(out)), "remc_get_temperatures")
        return out

    def counters(self):
        arr = (nat.remc_counters * self.n_replicas)()
        self._chk(self.lib.remc_get_counters(self.ctx, arr), "remc_get_counters")
        return np.array([[c.accept, c.accept_boltz, c.reject, c.steps] for c in arr], dtype=np.uint64)

    def ladder(self, sys_id):
        n = self.replica_sys.count(sys_id)
        out = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.remc_get_ladder(self.ctx, sys_id, nat.iptr(out)), "remc_get_ladder")
        return out

    def state(self, replica):
        sa = self.systems[self.replica_sys[replica]]
        n, m = sa.n_beads, sa.n_bodies
        b = [np.zeros(n, dtype=np.float32) for _ in range(6)]
        c = [np.zeros(m, dtype=np.float32) for _ in range(3)]
        q = np.zeros((m, 4), dtype=np.float64)
        self._chk(self.lib.remc_get_state(self.ctx, replica, *[nat.fptr(a) for a in b + c], nat.dptr(q)),
                  "remc_get_state")
        return dict(pos=np.stack(b[:3], 1), rel=np.stack(b[3:], 1), center=np.stack(c, 1), quat=q)

    def set_state(self, replica, pos=None, rel=None, center=None, quat=None):
        def cols(a, k):
            return [None] * k if a is None else [np.ascontiguousarray(a[:, i], dtype=np.float32) for i in range(k)]
        arrs = cols(pos, 3) + cols(rel, 3) + cols(center, 3)
        ptrs = [nat.fptr(a) if a is not None else None for a in arrs]
        qa = None if quat is None else np.ascontiguousarray(quat, dtype=np.float64).reshape(-1)
        self._chk(self.lib.remc_set_state(self.ctx, replica, *ptrs, nat.dptr(qa) if qa is not None else None),
                  "remc_set_state")

    # ------------------------------------------------------ checkpoint ----
    def set_energies(self, energies):
        e = np.ascontiguousarray(energies, dtype=np.float64)
        self._chk(self.lib.remc_set_energies(self.ctx, nat.dptr(e)), "remc_set_energies")

    def set_counters(self, counters):
        """counters: (n_replicas, 4) accept, accept_boltz, reject, RNG steps."""
        arr = (nat.remc_counters * self.n_replicas)()
        for r, row in enumerate(np.asarray(counters, dtype=np.uint64)):
            arr[r].accept, arr[r].accept_boltz, arr[r].reject, arr[r].steps = (int(v) for v in row)
        self._chk(self.lib.remc_set_counters(self.ctx, arr), "remc_set_counters")

    def set_ladder(self, sys_id, slot_replica):
        a = np.ascontiguousarray(slot_replica, dtype=np.int32)
        self._chk(self.lib.remc_set_ladder(self.ctx, sys_id, nat.iptr(a)), "remc_set_ladder")

    def exchange_state(self, sys_id):
        rd, sw, of = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        self._chk(self.lib.remc_get_exchange_state(self.ctx, sys_id, ctypes.byref(rd), ctypes.byref(of),
                                                   ctypes.byref(sw)), "remc_get_exchange_state")
        return rd.value, of.value, sw.value

    def set_exchange_state(self, sys_id, round_, offset, swaps):
        self._chk(self.lib.remc_set_exchange_state(self.ctx, sys_id, int(round_), int(offset), int(swaps)),
                  "remc_set_exchange_state")

    # -------------------------------------------------------- analysis ----
    def drms(self, sys_id, ref_pos):
        """DRMS of every replica of a system against ref_pos (n_beads, 3)."""
        ref = np.asarray(ref_pos, dtype=np.float32)
        cols = [np.ascontiguousarray(ref[:, k]) for k in range(3)]
        out = np.zeros(self.replica_sys.count(sys_id), dtype=np.float64)
        self._chk(self.lib.remc_drms(self.ctx, sys_id, *(nat.fptr(c) for c in cols), nat.dptr(out)), "remc_drms")
        return out

    def radial_distribution(self, sys_id, body_type, n_types, nbin, lo, hi, site_bead=None):
        """g(r) by body-type pair over every replica of a system (gr.py's
        normalisation): array (n_types, n_types, nbin), zero below the diagonal."""
        bt = np.ascontiguousarray(body_type, dtype=np.int32)
        site = None if site_bead is None else np.ascontiguousarray(site_bead, dtype=np.int32)
        out = np.zeros(n_types * n_types * nbin, dtype=np.float64)
        self._chk(self.lib.remc_radial_distribution(self.ctx, int(sys_id), nat.iptr(bt),
                                                    None if site is None else nat.iptr(site), int(n_types), int(nbin),
                                                    float(lo), float(hi), nat.dptr(out)), "remc_radial_distribution")
        return out.reshape(n_types, n_types, nbin)

    def centre_histogram(self, sys_id, body_type, n_types, nbin, lo, hi):
        """Counts (n_types, n_types, nbin) of body-centre distances over the system's replicas."""
        bt = np.ascontiguousarray(body_type, dtype=np.int32)
        out = np.zeros(n_types * n_types * nbin, dtype=np.uint64)
        self._chk(self.lib.remc_centre_histogram(self.ctx, sys_id, nat.iptr(bt), int(n_types), int(nbin), float(lo),
                                                 float(hi), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                  "remc_centre_histogram")
        return out.reshape(n_types, n_types, nbin)

    # ---------------------------------------------------------- timing ----
    MC, FULL, PLAIN, EXCHANGE, ADJOINT = range(5)

    def timing_enable(self, on=True):
        self._chk(self.lib.remc_timing_enable(self.ctx, 1 if on else 0), "remc_timing_enable")

    def timing(self, cls):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self.lib.remc_timing_get(self.ctx, cls, ctypes.byref(ms), ctypes.byref(n)), "remc_timing_get")
        return ms.value, n.value
