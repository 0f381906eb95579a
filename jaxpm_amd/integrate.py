"""Fixed-step kick-drift-kick leapfrog and its checkpointed adjoint.

The reference drives ``make_ode_fn`` (ode.py:122-147) with external
integrators (diffrax / odeint) and differentiates them with
``RecursiveCheckpointAdjoint`` (tests/test_gradients.py:27-72).  This module
fixes one integrator, a kick-drift-kick leapfrog in the scale factor over the
same right-hand side ``dx/da = v / (a^3 E)``, ``dv/da = 1.5 Om F(x) / (a^2 E)``:

    v += h/2 * dv/da(a, x);  x += h v / (am^3 E(am)), am = a + h/2;
    v += h/2 * dv/da(a + h, x)

with the force at ``a + h`` reused for the next step's first half kick, so a
run of S steps costs S + 1 force evaluations.  Its adjoint is written by hand
(:class:`NBody`): the backward pass restores (x, v) from checkpoints taken
every ``checkpoint_every`` steps, recomputes one segment at a time keeping
the segment's positions and force meshes, and walks it backwards with
``jpm_force_vjp_f32`` (J^T of the force) and ``jpm_axpy_f32`` (the drift).
"""
import numpy as np
import torch

from . import _lib, _ops
from . import growth as gr
from ._convert import to_device
from ._lib import call, ptr, size, stream


def kdk_schedule(cosmo, a_steps):
    """Host float64 coefficients of the leapfrog over ``a_steps``.

    Returns (k0, beta, gamma): first half kick ``v += k0 F(x0)``, then for
    step i ``x += beta[i] v`` and ``v += gamma[i] F(x)``.
    """
    a = np.asarray(a_steps, np.float64)
    if a.ndim != 1 or a.size < 2 or np.any(np.diff(a) == 0):
        raise ValueError("a_steps must be a 1-D sequence of at least two distinct scale factors")
    om = cosmo.Omega_m
    c = 1.5 * om / (a ** 2 * np.sqrt(gr.Esqr(cosmo, a)))
    h = np.diff(a)
    am = a[:-1] + 0.5 * h
    beta = h / (am ** 3 * np.sqrt(gr.Esqr(cosmo, am)))
    gamma = 0.5 * h * c[1:]
    gamma[:-1] += 0.5 * h[1:] * c[1:-1]
    k0 = 0.5 * h[0] * c[0]
    return float(k0), beta, gamma


class Stepper:
    """Device buffers and C-ABI calls of one mesh / particle configuration."""

    def __init__(self, mesh_shape, n, relative=False, deterministic=False, r_split=0.0, kflags=0, atomic=False,
                 sort_tuning=None, reorder=True):
        self.shape = tuple(int(s) for s in mesh_shape)
        self.nx, self.ny, self.nz = self.shape
        self.n = int(n)
        self.m = self.nx * self.ny * self.nz
        if relative and self.n != self.m:
            raise ValueError("relative mode needs one particle per mesh cell")
        self.r_split = float(r_split)
        pflags = (_lib.RELATIVE if relative else 0) | (_lib.DETERMINISTIC if deterministic else 0) | \
            (_lib.ATOMIC if atomic else 0)
        self.pflags = pflags
        # (chunk_code, slice_code[, rho_global]) of the tile sort and the
        # gathers, see jaxpm_amd.autotune
        self.sort_tuning = tuple(sort_tuning) if sort_tuning else (0, 0)
        self.flags = pflags | (kflags << 8) | _lib.tune_flags(*self.sort_tuning)
        self.dev = _lib.device()
        # optional HIP-event trace: name -> list of (start, end) events around
        # each launch on torch's current stream (set by the benchmark)
        self.trace = None
        kb = size("jpm_force_kick_workspace_size", self.nx, self.ny, self.nz, self.n, self.flags)
        pb = size("jpm_force_vjp_pot_workspace_size", self.nx, self.ny, self.nz, self.n, self.flags)
        # the tile-sorted pipeline (potential kept per evaluation, stencil in
        # LDS) when the mesh and flags allow it; otherwise three force meshes
        self.sorted = kb > 0 and pb > 0
        self.sort_bytes = size("jpm_bin_sort_size", self.nx, self.ny, self.nz, self.n, self.flags) if self.sorted else 0
        if self.sorted:
            wb = max(kb, pb)
        else:
            wb = max(size("jpm_force_field_workspace_size", self.nx, self.ny, self.nz, self.n, self.flags),
                     size("jpm_force_vjp_workspace_size", self.nx, self.ny, self.nz, self.n, self.flags))
        self.ws = torch.empty(wb, dtype=torch.uint8, device=self.dev)
        # NBody runs its time loop on the particles in the tile order of the
        # initial positions (absolute positions on the sorted pipeline): each
        # tile's particles then sit together in memory, so the gathers' reads
        # and writes by particle index fill whole cache lines.  Results are
        # the same bits either way (the paint sums in fixed point, the
        # gathers are per particle).
        self.reorder = bool(reorder) and self.sorted and not relative
        # max |cotangent| kept from one force VJP to the next (see vjp_field_drift)
        self.cot_max = torch.zeros(1, dtype=torch.float32, device=self.dev)
        # tile counts of the last drifted positions and of the ones before
        # (see field_kick_drift), absolute positions on the sorted pipeline
        self._tile_counts = None
        if self.sorted and not relative:
            k = size("jpm_tile_count", self.nx, self.ny, self.nz)
            self._tile_counts = [torch.empty(k, dtype=torch.int32, device=self.dev) for _ in range(2)]

    def _counts(self, chain):
        """(counts_in, counts_out) of the next drift and flips the pair."""
        tc = self._tile_counts
        if tc is None:
            return None, None
        cin = tc[0] if chain else None
        tc.reverse()
        return cin, tc[0]

    def counts_save(self, out):
        """Copy the per-tile counts of the latest drift's positions into
        ``out`` (a tensor from counts_like); no-op without tile counts."""
        if self._tile_counts is not None and out is not None:
            out.copy_(self._tile_counts[0])

    def counts_load(self, src):
        """Make ``src`` (from counts_save) the counts the next chained sort
        takes: its positions must be the ones saved then, unchanged."""
        if self._tile_counts is not None and src is not None:
            self._tile_counts[0].copy_(src)

    def counts_like(self):
        """A buffer for counts_save, or None (relative positions)."""
        return None if self._tile_counts is None else torch.empty_like(self._tile_counts[0])

    def tile_order(self, x):
        """order[k] = index of the particle at slot k of the tile sort of x."""
        return _ops.tile_order(x, self.shape, self.ws)

    def permute(self, order, *arrays, scatter=False):
        return _ops.permute_rows(order, *arrays, scatter=scatter)

    def new_phi(self):
        return torch.empty((3, *self.shape), dtype=torch.float32, device=self.dev)

    def new_field(self, keep_sort=False):
        """Buffer for what one force evaluation keeps for its adjoint: the
        potential and (``keep_sort``) the tile sort of the positions on the
        sorted pipeline, or the three force meshes."""
        if self.sorted:
            pot = torch.empty(self.shape, dtype=torch.float32, device=self.dev)
            srt = torch.empty(self.sort_bytes, dtype=torch.uint8, device=self.dev) if keep_sort else None
            return (pot, srt)
        return (self.new_phi(), None)

    def field_kick(self, x, v, gamma, field, chain=False):
        """v += gamma F(x); ``field`` (from new_field) receives what the
        adjoint of this evaluation needs.  ``v`` None: the force evaluation
        alone (its field, no kick).  ``chain`` as in field_kick_drift: x is
        the previous drift's output, whose tile counts replace the counting
        pass."""
        buf, srt = field
        if self.sorted:
            tc = self._tile_counts
            cin = tc[0] if chain and tc is not None else None
            self._timed("field_kick", "jpm_force_kick_f32", stream(), self.nx, self.ny, self.nz, ptr(x), self.n,
                        ptr(v), float(gamma), ptr(buf), self.r_split, self.flags, 0, ptr(self.ws),
                        self.ws.numel(), ptr(srt), 0 if srt is None else srt.numel(), ptr(cin))
        else:
            self.field(x, buf)
            if v is not None:
                self.kick(buf, x, v, gamma)

    def forces(self, x, out, field):
        """out = F(x) (the force field_kick would kick with); ``field``
        receives the evaluation's adjoint data as in field_kick."""
        buf, srt = field
        if self.sorted:
            self._timed("forces", "jpm_force_kick_f32", stream(), self.nx, self.ny, self.nz, ptr(x), self.n,
                        ptr(out), 1.0, ptr(buf), self.r_split, self.flags, 1, ptr(self.ws), self.ws.numel(),
                        ptr(srt), 0 if srt is None else srt.numel(), None)
        else:
            self.field(x, buf)
            out.zero_()
            self.kick(buf, x, out, 1.0)

    def field_kick_drift(self, x, v, gamma, beta, field, xout, xkeep=None, chain=False, vout=None):
        """v += gamma F(x), then the drift xout = x + beta v (also into
        ``xkeep`` when given); ``xout`` may be ``x``.  One launch sequence on
        the sorted pipeline (the drift rides on the kick's gather).  With
        absolute positions the kick also counts xout's particles per tile for
        the next sort; ``chain`` says x is the previous field_kick_drift's
        xout, unchanged since, whose counts then replace the counting pass.
        ``vout`` (not overlapping v): the kicked velocity goes there and v is
        only read."""
        buf, srt = field
        if self.sorted:
            cin, cout = self._counts(chain)
            self._timed("field_kick", "jpm_force_kick_drift_f32", stream(), self.nx, self.ny, self.nz, ptr(x), self.n,
                        ptr(v), float(gamma), float(beta), ptr(xout), ptr(xkeep), ptr(buf), self.r_split,
                        self.flags, ptr(self.ws), self.ws.numel(), ptr(srt), 0 if srt is None else srt.numel(),
                        ptr(cin), ptr(cout), ptr(vout))
        else:
            if vout is not None:
                vout.copy_(v)
                v = vout
            self.field_kick(x, v, gamma, field)
            if xout is not x:
                xout.copy_(x)
            self.axpy(beta, v, xout)
            if xkeep is not None:
                xkeep.copy_(xout)

    def vjp_field_drift(self, field, x, cot, scale, beta, gpos, chain=False):
        """gpos += scale J(x)^T cot, then the adjoint of the drift before the
        kick, cot += beta gpos.  The call leaves max |cot| of the updated
        cotangent in ``self.cot_max``; ``chain`` says the cotangent is the one
        the previous vjp_field_drift updated (unchanged since), whose kept
        maximum then bounds the cotangent paint without another pass."""
        buf, srt = field
        if self.sorted:
            self._timed("vjp", "jpm_force_vjp_pot_drift_f32", stream(), self.nx, self.ny, self.nz, ptr(buf), ptr(x),
                        self.n, ptr(cot), float(scale), float(beta), ptr(gpos), self.r_split, self.flags,
                        ptr(self.ws), self.ws.numel(), ptr(srt), 0 if srt is None else srt.numel(),
                        ptr(self.cot_max) if chain else None, ptr(self.cot_max))
        else:
            self.vjp_field(field, x, cot, scale, gpos)
            self.axpy(beta, gpos, cot)

    def vjp_field(self, field, x, cot, scale, gpos):
        """gpos += scale J(x)^T cot given the evaluation's ``field``."""
        buf, srt = field
        if self.sorted:
            self._timed("vjp", "jpm_force_vjp_pot_f32", stream(), self.nx, self.ny, self.nz, ptr(buf), ptr(x),
                        self.n, ptr(cot), float(scale), ptr(gpos), self.r_split, self.flags, ptr(self.ws),
                        self.ws.numel(), ptr(srt), 0 if srt is None else srt.numel())
        else:
            self.vjp(buf, x, cot, scale, gpos)

    def _timed(self, name, fn, *args):
        if self.trace is None:
            return call(fn, *args)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        call(fn, *args)
        e1.record()
        self.trace.setdefault(name, []).append((e0, e1))

    def _ensure_ws(self, name):
        need = size(name, self.nx, self.ny, self.nz, self.n, self.flags)
        if self.ws.numel() < need:
            self.ws = torch.empty(need, dtype=torch.uint8, device=self.dev)

    def field(self, x, phi):
        self._ensure_ws("jpm_force_field_workspace_size")
        self._timed("field", "jpm_force_field_f32", stream(), self.nx, self.ny, self.nz, ptr(x), self.n, None,
                    ptr(phi), self.r_split, 1.0, self.flags, ptr(self.ws), self.ws.numel())

    def kick(self, phi, x, v, gamma):
        self._timed("kick", "jpm_kick_f32", stream(), self.nx, self.ny, self.nz, ptr(phi), ptr(x), self.n,
                    float(gamma), ptr(v), self.pflags & _lib.RELATIVE)

    def axpy(self, a, x, y):
        self._timed("axpy", "jpm_axpy_f32", stream(), 3 * self.n, float(a), ptr(x), ptr(y))

    def vjp(self, phi, x, cot, scale, gpos):
        self._ensure_ws("jpm_force_vjp_workspace_size")
        self._timed("vjp", "jpm_force_vjp_f32", stream(), self.nx, self.ny, self.nz, ptr(phi), ptr(x), self.n,
                    ptr(cot), float(scale), ptr(gpos), self.r_split, self.flags, ptr(self.ws), self.ws.numel())


def schedule_tensor(sched):
    """(k0, beta, gamma) -> the float64 coefficient vector [k0, beta, gamma]
    that :class:`NBody` takes."""
    k0, beta, gamma = sched
    return torch.as_tensor(np.concatenate([[k0], np.asarray(beta, np.float64), np.asarray(gamma, np.float64)]),
                           dtype=torch.float64)


def _split(coeffs):
    c = coeffs.detach().to(torch.float64).cpu().numpy()
    if c.ndim != 1 or c.size < 3 or c.size % 2 == 0:
        raise ValueError("leapfrog coefficients must be [k0, beta(S), gamma(S)]")
    S = (c.size - 1) // 2
    return float(c[0]), c[1:1 + S], c[1 + S:]


def _run(stepper, x, v, k0, beta, gamma, start, stop, phi, keep=None, forces=False):
    """Advance steps [start, stop) in place.  With ``keep`` (a list of
    preallocated per-step buffers) each step's positions are copied out and
    its force evaluation writes into the step's field buffer; with
    ``forces`` the step also keeps the velocity it drifted with and the
    forces it kicked with (v += gamma F as forces-then-axpy, the same float
    operations as the fused kick)."""
    for i in range(start, stop):
        stepper.axpy(beta[i], v, x)
        if keep is not None:
            kp = keep[i - start]
            kp[0].copy_(x)
            phi = kp[1]
            if forces:
                kp[2].copy_(v)
                stepper.forces(x, kp[3], phi)
                stepper.axpy(gamma[i], kp[3], v)
                continue
        stepper.field_kick(x, v, gamma[i], phi)


def _run_fused(stepper, y, v, beta, gamma, start, stop, phi, keep=None, work=None, out=None, chain=False,
               kick_last=True, vwork=None, vout=None):
    """Kicks [start, stop) of the fused form.  On entry ``y`` holds y_start
    (the positions kick ``start`` evaluates its force at, i.e. after drift
    ``start``) and is only read; v is the velocity before that kick and is
    kicked in place.  Each kick carries the next drift,
    y_{i+1} = y_i + beta_{i+1} v, written straight into the buffer y_{i+1}
    lives in: with ``keep``, keep[i + 1 - start][0] (each step's positions
    and field stay there for the adjoint), otherwise ``work`` (in place
    there after the first drift).  The segment's last kick drifts into
    ``out`` when given and a drift follows (none after the final step), and
    drifts nowhere otherwise.  So checkpoints and kept positions are the
    buffers the drifts write, never copies.  Returns the buffer of the
    positions the segment's last kick evaluated.  ``chain``: y is the output
    of the stepper's previous drift (its tile counts are kept); every later
    kick of the segment follows a drift of its own.  Without ``kick_last`` a
    last kick that drifts nowhere is the force evaluation alone (an adjoint's
    recompute needs its field, not its velocity).  Velocities: in place on
    v, or with ``vwork`` v is only read, the kicks writing vwork (in place
    there after the first) and the segment's last drifting kick ``vout``
    when given.  Returns (positions of the last kick, velocity buffer)."""
    S = len(beta)
    cur, vc = y, v
    for i in range(start, stop):
        field = keep[i - start][1] if keep is not None else phi
        if i + 1 < S and (i + 1 < stop or out is not None):
            nxt = out if i + 1 == stop else (keep[i + 1 - start][0] if keep is not None else work)
            vn = vout if i + 1 == stop and vout is not None else (vwork if vwork is not None else vc)
            stepper.field_kick_drift(cur, vc, gamma[i], beta[i + 1], field, nxt, chain=chain or i > start,
                                     vout=None if vn is vc else vn)
            vc = vn
            if i + 1 < stop:
                cur = nxt
        else:
            stepper.field_kick(cur, vc if kick_last or i + 1 < stop else None, gamma[i], field,
                               chain=chain or i > start)
    return cur, vc


def _keep_buffers(st, like, count, forces, first_pos=True):
    """Per-step adjoint data of a segment: positions, the force evaluation's
    field (potential + kept tile sort) and, for coefficient cotangents, the
    drift velocity and the forces.  Without ``first_pos`` the first step's
    positions get no buffer (the fused sweeps read them from the segment's
    checkpoint)."""
    out = []
    for j in range(count):
        b = [torch.empty_like(like) if first_pos or j > 0 else None, st.new_field(keep_sort=True)]
        if forces:
            b += [torch.empty_like(like), torch.empty_like(like)]
        out.append(b)
    return out


class NBody(torch.autograd.Function):
    """(x0, v0, c) -> (xS, vS) by the leapfrog, with a hand-written adjoint.

    ``c`` is the float64 host vector [k0, beta(S), gamma(S)] of
    :func:`kdk_schedule`.  Its cotangent, when asked for, is
    dL/dk0 = <lambda_v, F(x0)>, dL/dbeta_i = <lambda_x, v_i>,
    dL/dgamma_i = <lambda_v, F(x_{i+1})>, each a float64 device dot product
    of the adjoint state with the trajectory.

    The forward keeps (x, v) at the start of every segment of
    ``checkpoint_every`` steps, and the adjoint data of the first half kick
    and of the last segment, which the backward would otherwise recompute
    first; every other segment is recomputed from its checkpoint."""

    @staticmethod
    def forward(ctx, x0, v0, coeffs, stepper, checkpoint_every):
        st = stepper
        k0, beta, gamma = _split(coeffs)
        S = len(beta)
        grad = any(ctx.needs_input_grad[:3])
        want_c = ctx.needs_input_grad[2]
        # the fused sweeps run on the particles in tile order (Stepper.reorder)
        order = st.tile_order(x0.detach()) if st.reorder and not want_c else None
        if order is not None:
            x, v = st.permute(order, x0.detach(), v0.detach())
        else:
            x = x0.detach().clone().contiguous()
            v = v0.detach().clone().contiguous()
        ctx.order = order
        phi0 = st.new_field(keep_sort=grad)
        f0 = None
        k = S if checkpoint_every is None else max(1, int(checkpoint_every))
        last = ((S - 1) // k) * k
        ctx.stepper, ctx.sched, ctx.k = st, (k0, beta, gamma), k
        ctx.c_meta = (coeffs.dtype, coeffs.device)
        ctx.fused = not want_c
        if ctx.fused:
            # every kick carries the drift that follows it, written into the
            # buffer of the positions it makes; checkpoints hold (y, v) just
            # before a kick, y being the buffer the drift wrote (no copy)
            y = torch.empty_like(x) if grad else x
            # checkpoint velocities are the buffers the kicks write, too
            vckpt = (lambda s: torch.empty_like(v) if grad and s != last else None)  # noqa: E731
            vc = vckpt(0)
            st.field_kick_drift(x, v, k0, beta[0], phi0, y, vout=vc)
            v = vc if vc is not None else v
            x0c = x if grad else None
            keep = _keep_buffers(st, x, min(k, S), False, first_pos=False) if grad else None
            work = torch.empty_like(x) if grad else x
            vwork = torch.empty_like(v) if grad and last > 0 else None
            phi = st.new_field()
            ckpts = {}
            for s0 in range(0, S, k):
                s1 = min(S, s0 + k)
                kept = s0 == last
                if grad and not kept:
                    # the checkpoint's positions, velocity and tile counts (those
                    # of the drift that wrote y, for the recompute's first sort)
                    c = st.counts_like()
                    st.counts_save(c)
                    ckpts[s0] = (y, v, c)
                elif grad:
                    # the kept last segment is never recomputed
                    ckpts[s0] = (y, None, None)
                out = (torch.empty_like(x) if grad else work) if s1 < S else None
                yl, v = _run_fused(st, y, v, beta, gamma, s0, s1, phi, keep if grad and kept else None, work=work,
                                   out=out, chain=True, vwork=None if kept else vwork,
                                   vout=vckpt(s1) if s1 < S else None)
                y = out
            ctx.ckpts = ckpts
            ctx.kept = (keep, phi0, x0c, last)
            if order is not None:
                return st.permute(order, yl, v, scatter=True)
            # kept positions are overwritten by the adjoint's recompute
            return (yl.clone() if grad else yl), v
        # with coefficient cotangents: separate forces and axpy, keeping the
        # forces and drift velocities the cotangents are dot products with
        f0 = torch.empty_like(x)
        st.forces(x, f0, phi0)
        st.axpy(k0, f0, v)
        ckpts = {0: (x.clone(), v.clone())}   # x0 and v after the first half kick
        keep = _keep_buffers(st, x, min(k, S), want_c) if grad else None
        phi = st.new_field()
        for s0 in range(0, S, k):
            if grad and s0 == last:
                _run(st, x, v, k0, beta, gamma, s0, min(S, s0 + k), phi, keep, forces=want_c)
            else:
                _run(st, x, v, k0, beta, gamma, s0, min(S, s0 + k), phi)
            if s0 + k < S:
                ckpts[s0 + k] = (x.clone(), v.clone())
        ctx.ckpts = ckpts
        ctx.kept = (keep, phi0, f0, last)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        if ctx.fused:
            return NBody._backward_fused(ctx, gx, gv)
        st = ctx.stepper
        k0, beta, gamma = ctx.sched
        S = len(beta)
        want_c = ctx.needs_input_grad[2]
        keep, phi0, f0, last = ctx.kept
        x0 = ctx.ckpts[0][0]
        lx = (gx if gx is not None else torch.zeros_like(x0)).detach().clone().contiguous()
        lv = (gv if gv is not None else torch.zeros_like(x0)).detach().clone().contiguous()
        dots = torch.zeros(1 + 2 * S, dtype=torch.float64, device=x0.device) if want_c else None
        phi = st.new_field()
        for s0 in reversed(range(0, S, ctx.k)):
            s1 = min(S, s0 + ctx.k)
            if s0 != last:
                xc, vc = ctx.ckpts[s0]
                x, v = xc.clone(), vc.clone()
                _run(st, x, v, k0, beta, gamma, s0, s1, phi, keep, forces=want_c)
            for i in reversed(range(s0, s1)):
                kp = keep[i - s0]
                if want_c:
                    _ops.dot_into(dots, 1 + S + i, lv, kp[3])     # gamma_i
                st.vjp_field(kp[1], kp[0], lv, gamma[i], lx)      # kick i: v += gamma F(x_{i+1})
                if want_c:
                    _ops.dot_into(dots, 1 + i, lx, kp[2])         # beta_i
                st.axpy(beta[i], lx, lv)                          # drift i: x += beta v
        # the first half kick, v += k0 F(x0), from its kept field
        if want_c:
            _ops.dot_into(dots, 0, lv, f0)
        st.vjp_field(phi0, x0, lv, k0, lx)
        ctx.ckpts = ctx.kept = None
        gc = dots.to(device=ctx.c_meta[1], dtype=ctx.c_meta[0]) if want_c else None
        return lx, lv, gc, None, None

    @staticmethod
    def _backward_fused(ctx, gx, gv):
        """The adjoint of the fused forward: per kick i, in reverse, the force
        VJP at y_i followed by the adjoint of drift i (lv += beta_i lx), one
        launch sequence; then the first half kick's VJP at x0."""
        st = ctx.stepper
        k0, beta, gamma = ctx.sched
        S = len(beta)
        keep, phi0, x0, last = ctx.kept
        like = ctx.ckpts[0][0]
        order = ctx.order
        gx = (gx if gx is not None else torch.zeros_like(like)).detach()
        gv = (gv if gv is not None else torch.zeros_like(like)).detach()
        if order is not None:
            lx, lv = st.permute(order, gx, gv)
        else:
            lx, lv = gx.clone().contiguous(), gv.clone().contiguous()
        phi = st.new_field()
        for s0 in reversed(range(0, S, ctx.k)):
            s1 = min(S, s0 + ctx.k)
            y, v, c = ctx.ckpts.pop(s0)
            if s0 != last:
                # each checkpoint is restored once: recompute in place, its
                # first sort from the saved tile counts
                st.counts_load(c)
                _run_fused(st, y, v, beta, gamma, s0, s1, phi, keep, kick_last=False, chain=c is not None)
            for i in reversed(range(s0, s1)):
                kp = keep[i - s0]
                st.vjp_field_drift(kp[1], y if i == s0 else kp[0], lv, gamma[i], beta[i], lx, chain=i < S - 1)
        st.vjp_field(phi0, x0, lv, k0, lx)
        ctx.ckpts = ctx.kept = ctx.order = None
        if order is not None:
            return (*st.permute(order, lx, lv, scatter=True), None, None, None)
        return lx, lv, None, None, None


def leapfrog_autodiff(x0, v0, coeffs, mesh_shape, paint_absolute_pos=True, trajectory=None):
    """The plain form of the same leapfrog, for comparison with
    :class:`NBody`: every step is a differentiable tensor expression
    (``pm_forces`` through the per-operation adjoints of ``_autograd``, kicks
    and drifts as tensor arithmetic) and autograd differentiates it over the
    stored trajectory -- the analogue of jax.grad through the reference's
    integrator with all intermediates kept (no checkpointing, no fused
    kernels).  ``trajectory``: a list that receives the (detached) positions
    of every force evaluation, the initial ones first."""
    from .pm import pm_forces
    k0, beta, gamma = _split(coeffs)

    def force(x):
        if trajectory is not None:
            trajectory.append(x.detach())
        return pm_forces(x, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos)

    x = x0
    v = v0 + k0 * force(x0)
    for i in range(len(beta)):
        x = x + float(beta[i]) * v
        v = v + float(gamma[i]) * force(x)
    return x, v


def gradient_gap(xa, xb, ga, gb, trajectory=None, top=100):
    """Where two float32 evaluations of the same leapfrog gradient differ.

    ``xa``, ``xb``: the final (absolute, grid-unit) positions each evaluation
    reached; ``ga``, ``gb``: its gradients with respect to the initial
    positions, all ``(N, 3)``; ``trajectory``: optionally, the positions of
    every force evaluation of evaluation a.  CIC weights are piecewise
    linear, so the forces' and the painted loss's derivatives jump where a
    particle crosses a cell face: two evaluations whose trajectories differ by
    float rounding put a particle that comes within that rounding of a face on
    different pieces, and its gradient then differs at O(1).  A particle is
    "at a face" when its final cells differ between the two evaluations, or a
    coordinate lies within ``tol`` of an integer at the end of either or at
    any force evaluation of the trajectory, with ``tol`` ten times the largest
    final position difference.  Returns the maximum difference relative to
    max |gb| over all particles and over the others, and how many of the
    ``top`` largest differences sit at a face."""
    xa, xb = xa.detach().double(), xb.detach().double()
    ga, gb = ga.detach().double(), gb.detach().double()
    dx = float((xa - xb).abs().max())
    tol = max(10.0 * dx, 1e-6)

    def face_dist(x):
        f = x - torch.floor(x)
        return torch.minimum(f, 1.0 - f)

    at_face = ((torch.floor(xa) != torch.floor(xb)).any(1) | (face_dist(xa) < tol).any(1) |
               (face_dist(xb) < tol).any(1))
    for xt in trajectory or ():
        at_face |= (face_dist(xt.detach().double().reshape(-1, 3)) < tol).any(1)
    scale = float(gb.abs().max())
    diff = (ga - gb).abs().amax(1)
    k = min(int(top), diff.numel())
    idx = torch.topk(diff, k).indices
    away = diff[~at_face]
    return {"max_rel": float(diff.max()) / scale,
            "max_rel_away_from_faces": (float(away.max()) / scale) if away.numel() else 0.0,
            "rms_rel_away_from_faces": (float(((ga - gb)[~at_face] ** 2).mean().sqrt()) /
                                        float((gb ** 2).mean().sqrt())) if away.numel() else 0.0,
            "final_position_max_diff_cells": dx, "face_tol_cells": tol,
            "particles_at_faces": int(at_face.sum()), "particles": int(diff.numel()),
            f"top{k}_at_faces": int(at_face[idx].sum())}


def leapfrog(pos, vel, a_steps, cosmo, mesh_shape, paint_absolute_pos=True, deterministic=False, r_split=0.0,
             checkpoint_every=None):
    """Integrate ``(pos, vel)`` over ``a_steps`` (absolute grid-unit positions,
    or displacements with ``paint_absolute_pos=False``).  Differentiable in
    both inputs through :class:`NBody`."""
    x0 = to_device(pos).reshape(-1, 3)
    v0 = to_device(vel).reshape(-1, 3)
    st = Stepper(mesh_shape, x0.shape[0], relative=not paint_absolute_pos, deterministic=deterministic,
                 r_split=r_split)
    coeffs = schedule_tensor(kdk_schedule(cosmo, a_steps))
    x, v = NBody.apply(x0, v0, coeffs, st, checkpoint_every)
    shape = to_device(pos).shape
    return x.reshape(shape), v.reshape(shape)


# ------------------------------------------------------------------ batches

class BatchStepper:
    """A batch of simulations of different meshes and particle counts, stepped
    by one launch per stage (``jpm_batch_*``).  Simulations of one mesh shape
    must be contiguous; arrays are the concatenation of the simulations in
    order: particles [N, 3], meshes [M]."""

    def __init__(self, mesh_shapes, counts, relative=False, r_split=0.0, kflags=0, sort_tuning=None, reorder=True):
        import ctypes
        self.shapes = [tuple(int(s) for s in sh) for sh in mesh_shapes]
        self.counts = [int(c) for c in counts]
        if len(self.shapes) != len(self.counts) or not self.shapes:
            raise ValueError("one particle count per mesh shape")
        self.B = len(self.shapes)
        self.relative = bool(relative)
        self.r_split = float(r_split)
        self.kflags = int(kflags)
        self.cells = [a * b * c for a, b, c in self.shapes]
        self.N = sum(self.counts)
        self.M = sum(self.cells)
        self.part0 = np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)
        self.cell0 = np.concatenate([[0], np.cumsum(self.cells)[:-1]]).astype(np.int64)
        self.dev = _lib.device()
        dims = (ctypes.c_int * (3 * self.B))(*[d for sh in self.shapes for d in sh])
        cnts = (ctypes.c_int64 * self.B)(*self.counts)
        h = ctypes.c_void_p()
        _lib.load()
        self.sort_tuning = tuple(sort_tuning) if sort_tuning else (0, 0)
        call("jpm_batch_create", self.B, dims, cnts,
             (_lib.RELATIVE if relative else 0) | _lib.tune_flags(*self.sort_tuning), ctypes.byref(h))
        self.handle = h
        self._destroy = _lib.load().jpm_batch_destroy
        self.sort_bytes = size("jpm_batch_sort_size", h)
        wb = max(size("jpm_batch_force_kick_workspace_size", h), size("jpm_batch_force_vjp_workspace_size", h))
        self.ws = torch.empty(wb, dtype=torch.uint8, device=self.dev)
        self.dot_ws = torch.empty(size("jpm_batch_dot_workspace_size", h), dtype=torch.uint8, device=self.dev)
        self.trace = None
        # BatchLeapfrog runs its sweeps on the particles in tile order, as
        # NBody does (Stepper.reorder); absolute positions only
        self.reorder = bool(reorder) and not self.relative
        self.cot_max = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self._tile_counts = None
        if not self.relative:
            k = size("jpm_batch_tile_count", h)
            self._tile_counts = [torch.empty(k, dtype=torch.int32, device=self.dev) for _ in range(2)]

    def tile_order(self, x, order):
        """order[k] = index of the particle at slot k of the batch's tile
        sort of x; each simulation's slots stay within its own range."""
        ws = self.ws if self.ws.numel() >= self.sort_bytes else \
            torch.empty(self.sort_bytes, dtype=torch.uint8, device=self.dev)
        call("jpm_batch_tile_order_u32", stream(), self.handle, ptr(x), ptr(order), ptr(ws), ws.numel())
        return order

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._destroy(h)
            self.handle = None

    def new_field(self, keep_sort=False):
        pot = torch.empty(self.M, dtype=torch.float32, device=self.dev)
        srt = torch.empty(self.sort_bytes, dtype=torch.uint8, device=self.dev) if keep_sort else None
        return (pot, srt)

    counts_save = Stepper.counts_save
    counts_load = Stepper.counts_load
    counts_like = Stepper.counts_like

    def _kick(self, x, v, gamma, field, write, chain=False):
        buf, srt = field
        tc = self._tile_counts
        cin = tc[0] if chain and tc is not None else None
        call("jpm_batch_force_kick_f32", stream(), self.handle, ptr(x), ptr(v), ptr(gamma), ptr(buf), self.r_split,
             self.kflags, int(write), ptr(self.ws), self.ws.numel(), ptr(srt), 0 if srt is None else srt.numel(),
             ptr(cin))

    def field_kick(self, x, v, gamma, field, chain=False):
        """v_b += gamma[b] F_b(x_b) for every simulation; gamma a float32
        device array [B].  ``v`` None: the force evaluations alone.
        ``chain`` as in field_kick_drift."""
        self._kick(x, v, gamma, field, False, chain)

    def forces(self, x, out, field, ones):
        """out = F(x) (``ones`` a float32 device array of 1.0, [B])."""
        self._kick(x, out, ones, field, True)

    def field_kick_drift(self, x, v, gamma, beta, field, xout, xkeep=None, chain=False, vout=None):
        """v_b += gamma[b] F_b(x_b), then xout_b = x_b + beta[b] v_b (also
        into ``xkeep``); gamma, beta float32 device arrays [B]; ``chain`` and
        ``vout`` as in Stepper.field_kick_drift."""
        buf, srt = field
        cin, cout = Stepper._counts(self, chain)
        call("jpm_batch_force_kick_drift_f32", stream(), self.handle, ptr(x), ptr(v), ptr(gamma), ptr(beta),
             ptr(xout), ptr(xkeep), ptr(buf), self.r_split, self.kflags, ptr(self.ws), self.ws.numel(), ptr(srt),
             0 if srt is None else srt.numel(), ptr(cin), ptr(cout), ptr(vout))

    def vjp_field_drift(self, field, x, cot, scale, pot_scale, beta, gpos, chain=False):
        """vjp_field, then the drift adjoint cot_b += beta[b] gpos_b; ``chain``
        and ``self.cot_max`` as in Stepper.vjp_field_drift."""
        buf, srt = field
        call("jpm_batch_force_vjp_drift_f32", stream(), self.handle, ptr(buf), ptr(x), ptr(cot), ptr(scale),
             ptr(pot_scale), ptr(beta), ptr(gpos), self.r_split, self.kflags, ptr(self.ws), self.ws.numel(),
             ptr(srt), 0 if srt is None else srt.numel(), ptr(self.cot_max) if chain else None, ptr(self.cot_max))

    def vjp_field(self, field, x, cot, scale, pot_scale, gpos):
        buf, srt = field
        call("jpm_batch_force_vjp_f32", stream(), self.handle, ptr(buf), ptr(x), ptr(cot), ptr(scale),
             ptr(pot_scale), ptr(gpos), self.r_split, self.kflags, ptr(self.ws), self.ws.numel(), ptr(srt),
             0 if srt is None else srt.numel())

    def axpy(self, a, x, y):
        call("jpm_batch_axpy_f32", stream(), self.handle, ptr(a), ptr(x), ptr(y))

    def dot_into(self, out, x, y):
        """out[:] (float64 device, [B]) = per-simulation <x_b, y_b>."""
        call("jpm_batch_dot_f32", stream(), self.handle, ptr(x), ptr(y), ptr(out), ptr(self.dot_ws),
             self.dot_ws.numel())

    def _lpt_ws(self):
        if getattr(self, "lpt_ws", None) is None:
            self.lpt_ws = torch.empty(size("jpm_batch_lpt_workspace_size", self.handle), dtype=torch.uint8,
                                      device=self.dev)
        return self.lpt_ws

    def linear_field(self, noise, boxes, ktab, pktabs):
        """Every simulation's Gaussian field from its noise (concatenated
        meshes [M]); boxes [B, 3], ktab [ntab], pktabs [B, ntab] on the device."""
        ws = self._lpt_ws()
        out = torch.empty_like(noise)
        call("jpm_batch_linear_field_f32", stream(), self.handle, ptr(noise), ptr(out), ptr(boxes), ptr(ktab),
             ptr(pktabs), ktab.numel(), ptr(ws), ws.numel())
        return out

    def lattice_forces(self, field):
        """[N, 3] first-order LPT forces at each simulation's lattice."""
        ws = self._lpt_ws()
        psi = torch.empty((self.N, 3), dtype=torch.float32, device=self.dev)
        call("jpm_batch_lattice_forces_f32", stream(), self.handle, ptr(field), ptr(psi), ptr(ws), ws.numel())
        return psi

    def lattice_forces_vjp(self, gpsi):
        ws = self._lpt_ws()
        g = torch.empty(self.M, dtype=torch.float32, device=self.dev)
        call("jpm_batch_lattice_forces_vjp_f32", stream(), self.handle, ptr(gpsi), ptr(g), ptr(ws), ws.numel())
        return g

    def paint(self, pos):
        """Concatenated CIC densities [M] of every simulation's particles."""
        mesh = torch.empty(self.M, dtype=torch.float32, device=self.dev)
        call("jpm_batch_paint_f32", stream(), self.handle, ptr(pos), ptr(mesh), ptr(self.ws), self.ws.numel())
        return mesh

    def paint_vjp(self, pos, gmesh):
        gpos = torch.empty((self.N, 3), dtype=torch.float32, device=self.dev)
        call("jpm_batch_paint_vjp_f32", stream(), self.handle, ptr(pos), ptr(gmesh), ptr(gpos))
        return gpos


class BatchSchedule:
    """Per-simulation leapfrog coefficients as the float32 device arrays the
    batched kernels read: k0 [B], beta [S, B], gamma [S, B], and for the
    adjoint the factors the potential pass applies, float32(gamma / M_b)
    and float32(k0 / M_b).  ``update`` rewrites them in place, so captured
    graphs see new values."""

    def __init__(self, coeffs, cells, dev):
        c = self._check(coeffs)
        B, S = c.shape[0], (c.shape[1] - 1) // 2
        self.B, self.S, self.cells, self.dev = B, S, np.asarray(cells, np.float64), dev
        z = lambda *sh: torch.zeros(sh, dtype=torch.float32, device=dev)  # noqa: E731
        self.k0, self.beta, self.gamma = z(B), z(S, B), z(S, B)
        self.gamma_pot, self.k0_pot = z(S, B), z(B)
        self.zeros = z(B)
        self.ones = torch.ones(B, dtype=torch.float32, device=dev)
        self.update(coeffs)

    @staticmethod
    def _check(coeffs):
        c = coeffs.detach().to(torch.float64).cpu().numpy() if isinstance(coeffs, torch.Tensor) else \
            np.asarray(coeffs, np.float64)
        if c.ndim != 2 or c.shape[1] < 3 or c.shape[1] % 2 == 0:
            raise ValueError("batched leapfrog coefficients must be [B, 1 + 2 S]")
        return c

    def update(self, coeffs):
        c = self._check(coeffs)
        if c.shape != (self.B, 1 + 2 * self.S):
            raise ValueError(f"coefficients of shape {c.shape}, expected {(self.B, 1 + 2 * self.S)}")
        S, m = self.S, self.cells
        f32 = c.astype(np.float32)
        host = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32))  # noqa: E731
        self.k0.copy_(host(f32[:, 0]))
        self.beta.copy_(host(f32[:, 1:1 + S].T))
        self.gamma.copy_(host(f32[:, 1 + S:].T))
        self.gamma_pot.copy_(host((f32[:, 1 + S:].astype(np.float64) / m[:, None]).astype(np.float32).T))
        self.k0_pot.copy_(host((f32[:, 0].astype(np.float64) / m).astype(np.float32)))


def _run_batch(st, sc, x, v, start, stop, phi, keep=None, forces=False):
    for i in range(start, stop):
        st.axpy(sc.beta[i], v, x)
        if keep is not None:
            kp = keep[i - start]
            kp[0].copy_(x)
            phi = kp[1]
            if forces:
                kp[2].copy_(v)
                st.forces(x, kp[3], phi, sc.ones)
                st.axpy(sc.gamma[i], kp[3], v)
                continue
        st.field_kick(x, v, sc.gamma[i], phi)


class BatchLeapfrog:
    """The batched leapfrog and its checkpointed adjoint over static device
    buffers.  ``forward`` / ``backward`` run the sweeps eagerly or, with
    ``graph=True``, replay two HIP graphs captured once (the time loop is a
    fixed sequence of stream-ordered launches: no host synchronisation, no
    allocation, no memset).  The adjoint is that of :class:`NBody`, per
    simulation; with ``coeff_grads`` the backward also forms the float64
    cotangents of every simulation's [k0, beta, gamma]."""

    def __init__(self, stepper, coeffs, checkpoint_every=None, coeff_grads=False, graph=False):
        st = self.st = stepper
        self.sc = BatchSchedule(coeffs, st.cells, st.dev)
        S = self.S = self.sc.S
        self.k = S if checkpoint_every is None else max(1, int(checkpoint_every))
        self.want_c = bool(coeff_grads)
        e = lambda: torch.empty((st.N, 3), dtype=torch.float32, device=st.dev)  # noqa: E731
        self.x0, self.x, self.v = e(), e(), e()
        self.last = ((S - 1) // self.k) * self.k
        # the fused sweeps keep no velocity for the kept last segment and
        # read each segment's first positions from its checkpoint
        fused = not self.want_c
        self.ckpts = {s0: (e(), e() if not (fused and s0 == self.last) else None) for s0 in range(0, S, self.k)}
        # the fused sweeps' checkpoint tile counts, for the recompute's first sort
        self.ckpt_counts = {s0: st.counts_like() for s0 in self.ckpts if fused and s0 != self.last}
        self.phi = st.new_field()
        self.phi0 = st.new_field(keep_sort=True)
        self.keep = _keep_buffers(st, self.x, min(self.k, S), self.want_c, first_pos=not fused)
        self.lx, self.lv = e(), e()
        # the fused forward's velocity work buffer (checkpoints are read only)
        self.vw = e() if fused and self.last > 0 else None
        self.vfinal = self.v
        if self.want_c:
            self.v0, self.xs, self.vs, self.f0 = e(), e(), e(), e()
        self.xfinal = self.x
        self._fresh = False
        self.dots = torch.zeros((1 + 2 * S, st.B), dtype=torch.float64, device=st.dev)
        # the fused sweeps (no coefficient cotangents) run on the particles in
        # tile order; forward / backward permute in and out around them
        self.order = torch.empty(st.N, dtype=torch.int32, device=st.dev) if st.reorder and not self.want_c else None
        if self.order is not None:
            self.xo, self.vo, self.lxo, self.lvo = e(), e(), e(), e()
        self.graphs = None
        if graph:
            self._capture()

    def _fwd(self):
        st, sc = self.st, self.sc
        x, v = self.x, self.v
        if not self.want_c:
            # every kick carries the drift that follows it, into the buffer
            # of the positions it makes (see _run_fused): from x0 into the
            # first checkpoint, then through x into the next; checkpoints
            # hold (y, v) just before a kick; v starts as the initial
            # velocity, which forward() writes there
            vck = lambda s: self.ckpts[s][1] if s < self.S else None  # noqa: E731
            st.field_kick_drift(self.x0, v, sc.k0, sc.beta[0], self.phi0, self.ckpts[0][0], vout=vck(0))
            y, v = self.ckpts[0][0], (vck(0) if vck(0) is not None else v)
            for s0 in range(0, self.S, self.k):
                s1 = min(self.S, s0 + self.k)
                kept = s0 == self.last
                if not kept:
                    st.counts_save(self.ckpt_counts[s0])
                out = self.ckpts[s1][0] if s1 < self.S else None
                self.xfinal, v = _run_fused(st, y, v, sc.beta, sc.gamma, s0, s1, self.phi,
                                            self.keep if kept else None, work=x, out=out, chain=True,
                                            vwork=None if kept else self.vw, vout=vck(s1))
                y = out
            self.vfinal = v
            return
        x.copy_(self.x0)
        v.copy_(self.v0)
        # the first half kick and the last segment keep their adjoint data
        st.forces(x, self.f0, self.phi0, sc.ones)
        st.axpy(sc.k0, self.f0, v)
        for s0 in range(0, self.S, self.k):
            if s0 == 0:
                self.ckpts[0][0].copy_(x)
                self.ckpts[0][1].copy_(v)
            if s0 == self.last:
                _run_batch(st, sc, x, v, s0, min(self.S, s0 + self.k), self.phi, self.keep, forces=self.want_c)
            else:
                _run_batch(st, sc, x, v, s0, min(self.S, s0 + self.k), self.phi)
            if s0 + self.k < self.S:
                self.ckpts[s0 + self.k][0].copy_(x)
                self.ckpts[s0 + self.k][1].copy_(v)

    def _bwd(self):
        st, sc, S = self.st, self.sc, self.S
        lx, lv, dots = self.lx, self.lv, self.dots
        if not self.want_c:
            for s0 in reversed(range(0, S, self.k)):
                s1 = min(S, s0 + self.k)
                y = self.ckpts[s0][0]
                if s0 != self.last:
                    # each checkpoint is read once per forward: recompute in
                    # place on its velocity
                    c = self.ckpt_counts[s0]
                    st.counts_load(c)
                    _run_fused(st, y, self.ckpts[s0][1], sc.beta, sc.gamma, s0, s1, self.phi, self.keep,
                               kick_last=False, chain=c is not None)
                for i in reversed(range(s0, s1)):
                    kp = self.keep[i - s0]
                    st.vjp_field_drift(kp[1], y if i == s0 else kp[0], lv, sc.gamma[i], sc.gamma_pot[i],
                                       sc.beta[i], lx, chain=i < S - 1)
            st.vjp_field(self.phi0, self.x0, lv, sc.k0, sc.k0_pot, lx)
            return
        if self.want_c:
            dots.zero_()
        for s0 in reversed(range(0, S, self.k)):
            s1 = min(S, s0 + self.k)
            if s0 != self.last:
                self.xs.copy_(self.ckpts[s0][0])
                self.vs.copy_(self.ckpts[s0][1])
                _run_batch(st, sc, self.xs, self.vs, s0, s1, self.phi, self.keep, forces=self.want_c)
            for i in reversed(range(s0, s1)):
                kp = self.keep[i - s0]
                if self.want_c:
                    st.dot_into(dots[1 + S + i], lv, kp[3])
                st.vjp_field(kp[1], kp[0], lv, sc.gamma[i], sc.gamma_pot[i], lx)
                if self.want_c:
                    st.dot_into(dots[1 + i], lx, kp[2])
                st.axpy(sc.beta[i], lx, lv)
        # the first half kick, v += k0 F(x0), from its kept field
        if self.want_c:
            st.dot_into(dots[0], lv, self.f0)
        st.vjp_field(self.phi0, self.ckpts[0][0], lv, sc.k0, sc.k0_pot, lx)

    def _capture(self):
        was = _lib.timing_enabled()
        _lib.timing(False)
        try:
            # one eager pass creates the FFT plans, wave-number tables and
            # kernel attributes outside the capture
            self._fwd()
            self._bwd()
            torch.cuda.synchronize()
            gf, gb = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
            with torch.cuda.graph(gf):
                self._fwd()
            with torch.cuda.graph(gb):
                self._bwd()
            torch.cuda.synchronize()
            self.graphs = (gf, gb)
        finally:
            _lib.timing(was)

    def forward(self, x0, v0, out=None):
        """Run the forward sweep from (x0, v0); returns (x, v) in ``out``
        when given, else in static buffers (valid until the next call)."""
        order = self.order
        vin = self.v0 if self.want_c else self.v
        if order is not None:
            # the order of this call's initial positions, outside the graph
            self.st.tile_order(x0, order)
            _ops.permute_rows(order, x0, v0, out=(self.x0, vin))
        else:
            self.x0.copy_(x0)
            vin.copy_(v0)
        if self.graphs:
            self.graphs[0].replay()
        else:
            self._fwd()
        self._fresh = True
        x, v = (self.x, self.v) if self.want_c else (self.xfinal, self.vfinal)
        if order is not None:
            return _ops.permute_rows(order, x, v, scatter=True, out=out if out is not None else (self.xo, self.vo))
        if out is not None:
            out[0].copy_(x)
            out[1].copy_(v)
            return out
        if x is not self.x:
            # the fused sweeps end in a kept buffer the adjoint overwrites
            self.x.copy_(x)
        return self.x, v

    def backward(self, gx, gv, out=None):
        """The adjoint of the last forward for output cotangents (gx, gv);
        returns (lx, lv) in ``out`` when given, else in static buffers, and
        the static ``dots``.  The fused sweeps recompute on their
        checkpoints in place, so each forward has one backward."""
        if not self._fresh:
            raise RuntimeError("BatchLeapfrog.backward needs a forward since the last backward")
        self._fresh = False
        order = self.order
        if order is not None:
            _ops.permute_rows(order, gx, gv, out=(self.lx, self.lv))
        else:
            self.lx.copy_(gx)
            self.lv.copy_(gv)
        if self.graphs:
            self.graphs[1].replay()
        else:
            self._bwd()
        if order is not None:
            res = _ops.permute_rows(order, self.lx, self.lv, scatter=True,
                                    out=out if out is not None else (self.lxo, self.lvo))
            return (*res, self.dots)
        if out is not None:
            out[0].copy_(self.lx)
            out[1].copy_(self.lv)
            return (*out, self.dots)
        return self.lx, self.lv, self.dots


class BatchNBody(torch.autograd.Function):
    """(x0, v0, c) -> (xS, vS) for a batch of simulations by the leapfrog,
    with the checkpointed hand-written adjoint of :class:`NBody`; ``c`` is
    the float64 host array [B, 1 + 2 S] of each simulation's [k0, beta,
    gamma] and receives per-simulation cotangents.  ``engine`` is a
    :class:`BatchStepper` (an eager :class:`BatchLeapfrog` is built per
    call) or a :class:`BatchLeapfrog`, possibly graph-captured, whose
    coefficients are set to ``c``."""

    @staticmethod
    def forward(ctx, x0, v0, coeffs, engine, checkpoint_every):
        if isinstance(engine, BatchStepper):
            engine = BatchLeapfrog(engine, coeffs, checkpoint_every, coeff_grads=ctx.needs_input_grad[2])
        else:
            engine.sc.update(coeffs)
            if ctx.needs_input_grad[2] and not engine.want_c:
                raise ValueError("this BatchLeapfrog was built without coefficient cotangents")
        x0, v0 = x0.detach().contiguous(), v0.detach().contiguous()
        x, v = engine.forward(x0, v0, out=(torch.empty_like(x0), torch.empty_like(v0)))
        ctx.engine = engine
        ctx.c_meta = (coeffs.dtype, coeffs.device)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        eng = ctx.engine
        zeros = None
        if gx is None or gv is None:
            zeros = torch.zeros_like(eng.x)
        gx = gx.contiguous() if gx is not None else zeros
        gv = gv.contiguous() if gv is not None else zeros
        lx, lv, dots = eng.backward(gx, gv, out=(torch.empty_like(gx), torch.empty_like(gv)))
        gc = None
        if ctx.needs_input_grad[2]:
            gc = dots.t().contiguous().to(device=ctx.c_meta[1], dtype=ctx.c_meta[0])
        ctx.engine = None
        return lx, lv, gc, None, None
