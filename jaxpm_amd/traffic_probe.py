"""Launch a fixed set of kernels for hardware-counter passes.

    python -m jaxpm_amd.traffic_probe [mesh] [chunk,slice,rho_global,xcd_slabs]

Runs, on a 256^3 (default) workload: a plain device copy of a known size
(the calibration launch), then the tile-sorted force evaluations of a 10-step
leapfrog from 2LPT initial conditions to the clustered state at a=1, on
the particles in the tile order of the initial positions as NBody runs
them, each kick carrying the next drift (`jpm_force_kick_drift_f32`, the
last one a plain kick), and 5 force VJPs with the drift adjoint
(`jpm_force_vjp_pot_drift_f32`) there: the launches of the leapfrog's
forward and adjoint sweeps.
Meant to be run under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`;
the per-dispatch counters of a kernel divided by those of the copy give its
HBM traffic in units of the copy's known byte count.
"""
import sys

import numpy as np
import torch


def main():
    from jaxpm_amd import _lib, integrate, pm
    from jaxpm_amd import cosmology as pc
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    tuning = tuple(int(c) for c in sys.argv[2].split(",")) if len(sys.argv) > 2 else (0, 0)
    dev = _lib.device()
    shape = (n, n, n)
    N = n ** 3
    # calibration: copy of 4N bytes (read 4N, write 4N)
    a = torch.randn(N, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    k, pk = pc.pk_table(cosmo)
    gen = torch.Generator(device=dev).manual_seed(2)
    with torch.no_grad():
        ic = pm.linear_field(shape, (4.0 * n,) * 3, pm.TabulatedPk(k, pk), torch.randn(shape, generator=gen, device=dev))
        lattice = torch.stack(torch.meshgrid(*[torch.arange(n, device=dev, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
        dx, p, _ = pm.lpt(cosmo, ic, lattice, a=0.1, order=2)
        x = (lattice + dx).reshape(-1, 3).contiguous()
        v = p.reshape(-1, 3).contiguous()
        st = integrate.Stepper(shape, N, sort_tuning=tuning)
        if st.reorder:
            x, v = st.permute(st.tile_order(x), x, v)
        steps = np.linspace(0.1, 1.0, 11)
        k0, beta, gamma = integrate.kdk_schedule(cosmo, steps)
        field = st.new_field(keep_sort=True)
        S = len(beta)
        st.field_kick_drift(x, v, k0, beta[0], field, x)
        for i in range(S):
            if i + 1 < S:
                st.field_kick_drift(x, v, gamma[i], beta[i + 1], field, x)
            else:
                st.field_kick(x, v, gamma[i], field)
        # force VJPs at the final state, reusing its kept tile sort
        cot = torch.randn((N, 3), generator=gen, device=dev)
        gpos = torch.zeros_like(x)
        for _ in range(5):
            st.vjp_field_drift(field, x, cot, float(gamma[-1]), float(beta[-1]), gpos)
    torch.cuda.synchronize()
    print(f"traffic probe done: copy of {4 * N} bytes, {len(beta) + 1} force evaluations and 5 force VJPs at "
          f"{n}^3, sort tuning {tuning}", flush=True)


if __name__ == "__main__":
    main()
