"""Autotuner for the tile sort and the gathers: picks the chunk width
(particles per accumulation / gather workgroup), the slice width (particles
per counting / scatter workgroup) and how tile workgroups are placed on the
XCDs (round-robin or one contiguous run per XCD) by timing one
force evaluation with kick and one force VJP on the caller's own particle
state (in the tile order the leapfrog runs it in), for every candidate.

Results are cached per (device, mesh, particle count, mode) in memory and in
``$XDG_CACHE_HOME/jaxpm_amd/sort_tuning.json`` (``~/.cache`` by default), so
the search runs once per configuration and machine.  The choice changes speed
only: every candidate computes the same forces up to float rounding.
"""
import json
import os

import torch

from . import _lib, _ops, integrate

CHUNK_CODES = (2, 3, 4, 5)      # 2048, 4096, 8192, 16384 particles
SLICE_CODES = (1, 2, 3, 4, 5)   # 4096, 8192, 16384, 32768, 65536 particles
# Where the force-VJP gather reads the density adjoint: only the LDS window
# (0) is searched.  With 512-thread workgroups both forms keep 4 waves per
# SIMD resident and the window moves a fifth fewer bytes; the cache form (1)
# wins on nearly uniform states (by ~4% at 256^3 2LPT initial conditions)
# and loses on clustered ones (by ~13% at a=1), where a leapfrog spends most
# of its steps, and the search only sees the initial state.
RHO_MODES = (0,)
SLAB_MODES = (0, 1)             # tiles dealt round-robin to the XCDs / in runs

_memo = {}


def _cache_path():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "jaxpm_amd", "sort_tuning.json")


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def _time_ms(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _key(kind, shapes, counts, relative):
    dev = torch.cuda.get_device_name(_lib.device())
    return f"{dev}|{kind}|{shapes}|{counts}|rel={int(bool(relative))}"


def _candidates():
    return [(c, s, rg, xs) for c in CHUNK_CODES for s in SLICE_CODES for rg in RHO_MODES for xs in SLAB_MODES]


def _cached(key, use_cache):
    if not use_cache:
        return None
    hit = _memo.get(key) or _load().get(key)
    if hit and len(hit.get("best", ())) == 4:
        _memo[key] = hit
        return tuple(hit["best"])
    return None


# Candidates within this factor of the fastest count as tied; ties go to the
# largest chunk width.  The search times the initial state, where tiles hold
# about their mean count and chunk widths barely matter; as the particles
# cluster, heavy tiles split into more chunks of a narrow width, each
# reloading its windows (4096 against 16384 particles: 234 against 270
# batch48 simulations/s), while timing noise alone can pick the narrow one.
TIE = 1.03
# Among those, slices within this factor of the fastest count as tied too,
# and ties go to the narrowest slice: on the 256^3 initial state slices of
# 4096 and 8192 particles time within 0.1% of each other, while over the
# whole 20-step leapfrog the 4096-particle scatter (512-thread workgroups,
# three per CU) runs 0.097 against 0.109 ms, 1% of the step.
SLICE_TIE = 1.01


def _finish(key, timings):
    fastest = min(timings.values())
    tied = [t for t, ms in timings.items() if ms <= TIE * fastest]
    widest = max(t[0] for t in tied)
    cands = [t for t in tied if t[0] == widest]
    near = min(timings[t] for t in cands) * SLICE_TIE
    best = min((t for t in cands if timings[t] <= near), key=lambda t: (t[1], timings[t]))
    _memo[key] = {"best": list(best), "ms": timings[best]}
    _store(key, _memo[key])
    return best, timings


def tune_stepper(mesh_shape, pos, relative=False, reps=8, use_cache=True):
    """Best (chunk_code, slice_code, rho_global, xcd_slabs) for a single simulation at
    positions ``pos`` ([n, 3] device tensor); returns (codes, {codes: ms})."""
    shape = tuple(int(s) for s in mesh_shape)
    n = pos.reshape(-1, 3).shape[0]
    key = _key("single", shape, n, relative)
    hit = _cached(key, use_cache)
    if hit:
        return hit, {}
    x = pos.reshape(-1, 3).contiguous()
    probe = integrate.Stepper(shape, n, relative=relative)
    if not probe.sorted:
        return (0, 0, 0, 0), {}
    if probe.reorder:
        # the particle order NBody's sweeps run in
        x = probe.permute(probe.tile_order(x), x)
    del probe
    cot = torch.randn_like(x)
    timings = {}
    for tune in _candidates():
        st = integrate.Stepper(shape, n, relative=relative, sort_tuning=tune)
        field = st.new_field(keep_sort=True)
        v = torch.zeros_like(x)
        g = torch.zeros_like(x)

        def step():
            st.field_kick(x, v, 0.0, field)
            st.vjp_field(field, x, cot, 1.0, g)

        timings[tune] = _time_ms(step, reps)
        del st, field
    return _finish(key, timings)


def tune_batch(mesh_shapes, counts, pos, relative=False, reps=3, use_cache=True):
    """Best (chunk_code, slice_code, rho_global, xcd_slabs) for a batch (BatchStepper
    layout)."""
    shapes = [tuple(int(s) for s in sh) for sh in mesh_shapes]
    counts = [int(c) for c in counts]
    key = _key("batch", shapes, counts, relative)
    hit = _cached(key, use_cache)
    if hit:
        return hit, {}
    x = pos.reshape(-1, 3).contiguous()
    probe = integrate.BatchStepper(shapes, counts, relative=relative)
    if probe.reorder:
        # the particle order BatchLeapfrog's sweeps run in
        order = probe.tile_order(x, torch.empty(probe.N, dtype=torch.int32, device=x.device))
        x = _ops.permute_rows(order, x)
    del probe
    cot = torch.randn_like(x)
    timings = {}
    for tune in _candidates():
        st = integrate.BatchStepper(shapes, counts, relative=relative, sort_tuning=tune)
        field = st.new_field(keep_sort=True)
        zero = torch.zeros(st.B, dtype=torch.float32, device=x.device)
        one = torch.ones(st.B, dtype=torch.float32, device=x.device)
        v = torch.zeros_like(x)
        g = torch.zeros_like(x)

        def step():
            st.field_kick(x, v, zero, field)
            st.vjp_field(field, x, cot, one, one, g)

        timings[tune] = _time_ms(step, reps)
        del st, field
    return _finish(key, timings)
