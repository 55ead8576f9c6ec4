"""DM gradients of folded profiles through torch.autograd, and gradient DM refinement.

``FoldedProfile`` maps trial DMs to folded profiles of a device-resident baseband record (MS-OSM
coherent dedispersion + fold, msosm_gpu_batch.cpp:178-236 and fold_gpu.cpp:48-60, on the GPU). Its
backward recomputes the dedispersed series: each block is dedispersed with H and with dH/dDM and
2 Re(conj(y) dy/dDM) is folded with the forward bins, which gives d profile / d DM; nothing from the
forward pass besides the DMs is kept. The chain rule with the incoming profile gradient gives the
DM gradient, so any torch loss on the profiles works (``snr_loss`` is the one used for refinement).

M and the ring depth are fixed when the engine is made (for DMs up to its ``dm_max``); the delay
blocks and the chirp of every trial are designed at the trial's own DM on each call, so a profile is
the library's MS-OSM at exactly that DM. DMs pass the reference interface as float32 (one ulp is
7.6e-6 pc/cm^3 at DM 123). The delay blocks are piecewise constant in DM, so the gradient is that
of the profile between the DMs where a frequency bin changes block.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lib as _load_library

# torch bundles its own HIP runtime; it must start before this library's (which starts on its first
# HIP call), or torch finds no GPU. Starting torch's here is a no-op where it already runs.
if torch.cuda.is_available():
    torch.cuda.init()
_load_library()  # the C ABI library must be present (no CPU fallback)
from . import _core  # noqa: E402  (pybind11 module, linked against the library)

ProfileEngine = _core.ProfileEngine


class FoldedProfile(torch.autograd.Function):
    """profile[D][nbins] (float64, CPU) = fold(|MS-OSM(record; DM_d)|^2); differentiable in dm."""

    @staticmethod
    def forward(ctx, dm: torch.Tensor, engine, record: torch.Tensor, n_samples: int):
        engine.set_dms(dm.detach().cpu().numpy().astype(np.float32))
        prof, _, _ = engine.run(record.data_ptr(), int(n_samples), False)
        ctx.engine, ctx.record, ctx.n = engine, record, int(n_samples)
        ctx.save_for_backward(dm)
        return torch.from_numpy(prof)

    @staticmethod
    def backward(ctx, grad_profile: torch.Tensor):
        (dm,) = ctx.saved_tensors
        engine = ctx.engine
        engine.set_dms(dm.detach().cpu().numpy().astype(np.float32), True)
        _, dprof, _ = engine.run(ctx.record.data_ptr(), ctx.n, True)
        g = (grad_profile.detach().to(torch.float64).cpu() * torch.from_numpy(dprof)).sum(dim=1)
        return g.to(dm.dtype), None, None, None


def folded_profile(dm: torch.Tensor, engine, record: torch.Tensor, n_samples: int | None = None) -> torch.Tensor:
    n = int(record.shape[0]) if n_samples is None else int(n_samples)
    return FoldedProfile.apply(dm, engine, record, n)


def snr_loss(profile: torch.Tensor, noise_var: float) -> torch.Tensor:
    """Per-DM loss -sum_b (p_b - mean(p))^2 / noise_var (lower is better)."""
    c = profile - profile.mean(dim=1, keepdim=True)
    return -(c * c).sum(dim=1) / noise_var


def profile_noise_var(record: torch.Tensor, n_samples: int, nbins: int) -> float:
    """Variance of one profile bin under noise alone: var(|x|^2) * nbins / n (fixed per record)."""
    x = record[:n_samples]
    if x.dtype == torch.int8:  # interleaved I/Q bytes, shape (n, 2)
        p = x.to(torch.float32).pow(2).sum(dim=-1)
    else:
        p = x.abs().pow(2).to(torch.float32)
    return float(p.to(torch.float64).var()) * nbins / n_samples


def refine_dm(engine, record: torch.Tensor, dm_start, noise_var: float, n_samples: int | None = None,
              steps: int = 40, step0: float = 0.02, step_min: float = 1e-5, step_max: float = 0.2,
              bounds: tuple | None = None):
    """Gradient DM refinement of every trial at once with sign-based steps (torch.optim.Rprop).

    Each step runs one forward and one backward (the derivative pass). The smallest step stays above
    one float32 ulp of the DMs. Trials are kept inside ``bounds`` (default: [0, engine.dm_max]).
    Returns (refined DMs, final losses, number of steps taken)."""
    lo, hi = bounds if bounds is not None else (0.0, float(engine.dm_max))
    dm = torch.tensor(np.clip(np.asarray(dm_start, dtype=np.float64), lo, hi), requires_grad=True)
    opt = torch.optim.Rprop([dm], lr=step0, etas=(0.5, 1.2), step_sizes=(step_min, step_max))
    loss = None
    for it in range(steps):
        opt.zero_grad()
        loss = snr_loss(folded_profile(dm, engine, record, n_samples), noise_var)
        loss.sum().backward()
        opt.step()
        with torch.no_grad():
            dm.clamp_(lo, hi)
    with torch.no_grad():
        final = snr_loss(folded_profile(dm.detach(), engine, record, n_samples), noise_var)
    return dm.detach().numpy(), final.numpy(), steps


def grid_losses(record: torch.Tensor, n_samples: int, bw: float, f0: float, dms, noise_var: float, nbins: int,
                period: float, batch: int, dm_max: float, chunk: int = 1024) -> np.ndarray:
    """Loss of every DM of a grid (engines of at most chunk trials, each trial designed at its DM)."""
    dms = np.asarray(dms, dtype=np.float64)
    out = []
    for s0 in range(0, dms.size, chunk):
        eng = ProfileEngine(bw, f0, dms[s0:s0 + chunk].astype(np.float32), batch, nbins, period, dm_max=dm_max)
        prof, _, _ = eng.run(record.data_ptr(), int(n_samples), False)
        out.append(snr_loss(torch.from_numpy(prof), noise_var).numpy())
        del eng
    return np.concatenate(out)


def search_dm(record: torch.Tensor, n_samples: int, bw: float, f0: float, lo: float, hi: float, noise_var: float,
              nbins: int, period: float, batch: int, coarse_step: float, polish: int = 4, steps: int = 20):
    """DM search: a coarse grid over [lo, hi], then gradient refinement from its `polish` deepest local
    minima. The coarse step should be near the DM scale on which the loss decorrelates (the residual
    chirp at the band edge reaching ~1 rad), so that every basin of the loss holds a coarse point and
    the refinement only has to descend within one.

    Returns a dict: best DM and loss, the coarse grid's best, the refined start and end points, and
    the number of DM-record passes spent (coarse trials plus three per refinement step and trial)."""
    grid = np.arange(lo, hi + 0.5 * coarse_step, coarse_step)
    dm_max = float(np.float32(hi))
    gl = grid_losses(record, n_samples, bw, f0, grid, noise_var, nbins, period, batch, dm_max)
    # local minima of the coarse losses, deepest first
    left = np.concatenate([[np.inf], gl[:-1]])
    right = np.concatenate([gl[1:], [np.inf]])
    minima = np.nonzero((gl <= left) & (gl <= right))[0]
    minima = minima[np.argsort(gl[minima])][:polish]
    starts = grid[minima]
    eng = ProfileEngine(bw, f0, starts.astype(np.float32), batch, nbins, period, dm_max=dm_max)
    dm, loss, nsteps = refine_dm(eng, record, starts, noise_var, n_samples, steps=steps, step0=coarse_step / 4,
                                 step_min=1e-5, step_max=coarse_step / 2, bounds=(lo, hi))
    b = int(np.argmin(loss))
    return {"best_dm": float(dm[b]), "best_loss": float(loss[b]), "coarse_best_dm": float(grid[int(np.argmin(gl))]),
            "coarse_best_loss": float(gl.min()), "coarse_trials": int(grid.size), "starts": starts.tolist(),
            "refined": dm.tolist(), "refined_losses": loss.tolist(),
            "passes": int(grid.size + len(starts) * (3 * nsteps + 1))}
