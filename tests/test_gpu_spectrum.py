"""CFS spectrum phase of the GPU path (stock -> rho sweep -> spikes -> bins)
against reference runs with spectra on: the SIAM workload (Nkeep=256, 40
iterations, two spectral types) and the two-orbital Kanamori model (Nkeep=300,
20 iterations, four spectral types, spin-triplet ground state).  Compared:
binned spectral weights A(omega) of both types and signs, the sum rule,
Tr rho per shell, and the gauge-invariant content of every reduced-density
block (trace and spectrum)."""
import os

import numpy as np
import pytest

from nrgamd import models

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
BIN_TOL = 1e-8      # absolute, relative to the largest bin weight
RHO_TOL = 1e-10     # absolute, on rho entries (eigenvalues <= 1)

CASES = {
    # fixture -> (workload, overrides)
    "siam_nk256_it40": ("siam_nk256_it40", {}),
    "two_orbital_nk300_it20_spec": ("two_orbital_nk4000_it45", dict(nkeep=300, niter=20, spectrum=True)),
}


@pytest.mark.parametrize("split", [0, 8], ids=["fp64", "split8"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_cfs_spectrum(tmp_path, name, split):
    # split 8: the invariant, spectral, CFS and rho products run as
    # split-precision FP16 matrix-core products (8 slices); same tolerances
    from nrgamd import gpu
    from nrgamd.host import Host
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    d = str(tmp_path / name)
    base, ov = CASES[name]
    models.write_inputs(models.workload(base, **ov), d)
    h = Host(d, device=0)
    gpu.load().nrg_gpu_set_gemm_split(split)
    try:
        print(f"[{name}] forward", flush=True)
        h.run_forward()
        i = h.info()
        assert i["start_trunc"] == int(fx["start_trunc"])
        print(f"[{name}] spectrum phase", flush=True)
        h.spectrum(keep_history=True)
        neg, pos = h.bins()
        assert neg.shape == fx["bins_neg"].shape and pos.shape == fx["bins_pos"].shape
        peak = max(np.abs(fx["bins_neg"]).max(), np.abs(fx["bins_pos"]).max())
        err = max(np.abs(neg - fx["bins_neg"]).max(), np.abs(pos - fx["bins_pos"]).max())
        print("bins: max abs error", err, "peak", peak)
        assert err <= BIN_TOL * peak
        # sum rule of the complete Fock space: sum over both signs = 1 per type
        np.testing.assert_allclose(neg.sum(1) + pos.sum(1), 1.0, atol=1e-10)

        st, last = i["start_trunc"], i["n_iter"]
        tr = h.rho_traces()
        for it in range(st, last + 1):
            assert abs(tr[it] - 1.0) < 1e-12, (it, tr[it])

        fs, mp = fx["full_subspace"], fx["map_full"]
        nv = i["nv"]
        ref_tr, ref_eig = fx["rho_block_trace"], fx["rho_block_eigs"]
        eig_start = np.concatenate(([0], np.cumsum(fs[:, nv + 1])))
        worst = 0.0
        for it in range(st, last + 1):
            kb = fx[f"it{it}_kept_blk"]
            blocks = [b for b in gpu.rho_get(it, kb) if b.size]
            r0 = int(mp[it - st]) - 1
            assert len(blocks) == int((kb > 0).sum())
            for j, b in enumerate(blocks):
                row = r0 + j
                assert int(fs[row, nv + 1]) == b.shape[0]
                e = np.linalg.eigvalsh(0.5 * (b + b.T))
                er = ref_eig[eig_start[row]:eig_start[row + 1]]
                worst = max(worst, abs(np.trace(b) - ref_tr[row]), np.abs(e - er).max())
        print("rho: worst block trace/eigenvalue error", worst)
        assert worst <= RHO_TOL
    finally:
        gpu.load().nrg_gpu_set_gemm_split(0)
        h.close()
