"""Surrogate-assisted data assimilation (no counterpart in the reference: new
in this library).

A window of `DataAssimilation` maps, per member, the hyper-parameters of its
voxels and the BOLD state at the window start to the end-of-window hidden
state tail (act, s, q, v, f_in, bold per voxel).  `WindowSurrogate` learns
that map with `MLPSurrogate`; `SurrogateDA` runs the window loop of
data_assimilation/DataAssimilation.py:402-453 and, in most windows, replaces
the network simulation by one surrogate evaluation.

Fallback policy: the simulator runs every `every`-th window, and also after
any simulated window whose surrogate NRMSE exceeds `nrmse_max`, where
NRMSE = ||BOLD_sur - BOLD_sim||_2 / (sqrt(n) * range(BOLD_sim)) over the
window's members and voxels.  Each simulated window adds its samples to the
training set and fine-tunes the surrogate.  Hyper-parameters are pushed into
the network only in simulated windows (the push is absolute: original value
times the current factor, so skipped windows lose nothing).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .surrogate import MLPSurrogate


NF = 9  # input features per voxel


def window_inputs(da):
    """[E, 9V]: per voxel the 4 hyper-parameter logits of its hidden state,
    the BOLD state (s, q, v, f_in) at the window start and the firing rate
    (act) of the previous window (network activity carries over)."""
    E = da.ensemble_number
    Vb = da._num_voxel_pblk
    idx = da._index_da_voxel_pblk.numpy()
    hp = da._hp_log.reshape(E, -1)[:, da.for_hidden_state].reshape(E, len(idx), -1).double()
    nv = da._da_voxels
    st = np.stack([da.block_model.nb.bold_state(w)[:nv] for w in ("s", "q", "v", "f_in")], axis=1)
    st = torch.from_numpy(st.reshape(E, Vb, 4)[:, idx])
    if da._hidden_state is not None:
        act = da._hidden_state[:, :, -6:-5].double().cpu()
    else:
        act = torch.zeros((E, len(idx), 1), dtype=torch.float64)
    return torch.cat((hp, st, act), dim=2).reshape(E, -1)


# default_params.py:11-12 (the BOLD output constants)
_V0, _E0 = 0.02, 0.8


def window_base(X, V, use_act=True):
    """What a window would end at if nothing changed: per voxel (the previous
    act (0 without use_act), the start s, q, v, f_in, and the BOLD signal of
    the start q, v as models/bold_model_pytorch.py:76-77 computes it).
    X: [E, V*NF]."""
    x = torch.as_tensor(X, dtype=torch.float64).reshape(X.shape[0], V, NF)
    s, q, v, f, act = x[..., 4], x[..., 5], x[..., 6], x[..., 7], x[..., 8]
    if not use_act:
        act = torch.zeros_like(act)
    bold = _V0 * (7 * _E0 * (1 - q) + 2 * (1 - q / v) + (2 * _E0 - 0.2) * (1 - v))
    return torch.stack((act, s, q, v, f, bold), dim=-1).reshape(X.shape[0], V * 6)


def nrmse(sur, sim, scale=None):
    """RMSE of sur against sim over all entries, divided by `scale` (default:
    the range of sim)."""
    sur = np.asarray(sur, np.float64)
    sim = np.asarray(sim, np.float64)
    rng = float(sim.max() - sim.min()) if scale is None else float(scale)
    return float(np.linalg.norm(sur - sim) / np.sqrt(sim.size) / max(rng, 1e-300))


class WindowSurrogate:
    """Window map window_inputs -> hidden-state tail [E, V, 6], learned as the
    change over the window (tail - window_base), so that the slow
    hemodynamic state needs no absolute precision from the network.

    per_voxel=True: one network shared by all voxels; a row is a voxel's own
    NF features and its member's mean of them (voxels of a member are
    exchangeable: every neuron of a member draws its inputs from the whole
    member), and each window gives E*V rows.  per_voxel=False: one row per
    member, all V voxels at once (V*NF inputs, V*6 outputs)."""

    def __init__(self, voxels, hidden=(512, 512, 512), max_batch=None, lr=1e-3, seed=0, device=0, per_voxel=True,
                 members=0, voxel_ids=False, use_act=True):
        self.V = int(voxels)
        self.per_voxel = bool(per_voxel)
        # use_act=False: the previous window's rate is left out of the input (the
        # DA's act is the last step's rate, mostly noise; after a surrogate window
        # it would be the surrogate's own smooth prediction instead)
        self.use_act = bool(use_act)
        self.nf = NF if self.use_act else NF - 1
        # members > 0 (per-member form): a one-hot of the member is part of the input, so the
        # network can learn what is particular to each member's own synapses and cells
        self.members = int(members) if not self.per_voxel else 0
        # voxel_ids (per-voxel form): a one-hot of the voxel, for voxels that are not
        # exchangeable (regions of different size and connectivity)
        self.voxel_ids = bool(voxel_ids) and self.per_voxel
        n_in, n_out = ((2 * self.nf + (self.V if self.voxel_ids else 0), 6) if self.per_voxel
                       else (self.V * self.nf + self.members, self.V * 6))
        mb = max_batch or (1024 if self.per_voxel else 256)
        self.mlp = MLPSurrogate(n_in, n_out, hidden=hidden, max_batch=mb, lr=lr, seed=seed, device=device)
        self.X = []
        self.Y = []

    def _rows(self, X, idx=None):
        X = torch.as_tensor(X, dtype=torch.float64).cpu()
        if not self.use_act:
            X = X.reshape(X.shape[0], self.V, NF)[:, :, :NF - 1].reshape(X.shape[0], -1)
        if not self.per_voxel:
            if self.members:
                idx = torch.arange(X.shape[0]) % self.members if idx is None else torch.as_tensor(idx)
                X = torch.cat((X, torch.nn.functional.one_hot(idx.long(), self.members).double()), dim=1)
            return X
        x = X.reshape(X.shape[0], self.V, self.nf)
        m = x.mean(dim=1, keepdim=True).expand(-1, self.V, -1)
        parts = [x, m]
        if self.voxel_ids:
            parts.append(torch.eye(self.V, dtype=torch.float64).expand(X.shape[0], -1, -1))
        return torch.cat(parts, dim=2).reshape(X.shape[0] * self.V, -1)

    def add(self, X, Y, idx=None, u=None):
        """Rows of X belong to members idx (default: row r is member r mod members);
        u (a window's mean BOLD input) is not used by this surrogate."""
        X = torch.as_tensor(X, dtype=torch.float64).cpu()
        Y = torch.as_tensor(Y, dtype=torch.float64).reshape(X.shape[0], -1).cpu()
        d = Y - window_base(X, self.V, self.use_act)
        self.X.append(self._rows(X, idx))
        self.Y.append(d.reshape(-1, 6) if self.per_voxel else d)

    def fit(self, epochs=None, batch=None, renormalize=True):
        X = torch.cat(self.X)
        Y = torch.cat(self.Y)
        if renormalize:
            self.mlp.set_normalization(X, Y)
        if epochs is None:
            epochs = 20 if self.per_voxel else 400
        return self.mlp.fit(X, Y, epochs=epochs, batch=batch)

    def finetune(self, epochs, replay=1.0, seed=0, max_rows=None):
        """Train on the last added window (at most max_rows of its rows, drawn
        at random) plus `replay` times as many rows drawn at random from the
        earlier ones (so the fit does not drift towards the latest window, at
        a small fraction of a pass over everything)."""
        Xn, Yn = self.X[-1], self.Y[-1]
        if max_rows is not None and Xn.shape[0] > max_rows:
            pick = torch.randperm(Xn.shape[0], generator=torch.Generator().manual_seed(seed + 1))[:max_rows]
            Xn, Yn = Xn.index_select(0, pick), Yn.index_select(0, pick)
        if len(self.X) > 1 and replay > 0:
            # distinct rows drawn from the earlier windows, gathered block by
            # block (no concatenation of everything stored so far)
            sizes = np.array([b.shape[0] for b in self.X[:-1]], np.int64)
            ends = np.cumsum(sizes)
            k = int(min(ends[-1], round(replay * Xn.shape[0])))
            rng = np.random.default_rng(seed)
            g = np.sort(rng.choice(ends[-1], size=k, replace=False)) if k < ends[-1] else np.arange(ends[-1])
            blk = np.searchsorted(ends, g, side="right")
            xs, ys = [Xn], [Yn]
            for b in np.unique(blk):
                loc = torch.from_numpy(g[blk == b] - (ends[b] - sizes[b]))
                xs.append(self.X[b].index_select(0, loc))
                ys.append(self.Y[b].index_select(0, loc))
            Xn, Yn = torch.cat(xs), torch.cat(ys)
        return self.mlp.fit(Xn, Yn, epochs=epochs, seed=seed)

    def residual_std(self, X, Y):
        """Per voxel and output the standard deviation of Y - predict(X) [V, 6]
        over the rows given: the spread of the simulator's windows around the
        surrogate's mean."""
        X = torch.as_tensor(X, dtype=torch.float64)
        Y = torch.as_tensor(Y, dtype=torch.float64).reshape(X.shape[0], self.V, 6)
        P = self.predict(X).double().cpu()
        return (Y - P).std(dim=0)

    def predict(self, X, idx=None):
        X = torch.as_tensor(X, dtype=torch.float64).cpu()
        d = self.mlp.predict(self._rows(X, idx)).double().cpu().reshape(X.shape[0], -1)
        return (d + window_base(X, self.V, self.use_act)).reshape(X.shape[0], self.V, 6).to(self.mlp.dev)


class SurrogateDA:
    """Every simulated window fine-tunes the surrogate on that window plus an
    equal-sized replay of earlier rows (`finetune_epochs` passes).

    Fallback: the window after a simulated one is simulated too when the
    surrogate's NRMSE on it exceeds nrmse_max and, with adapt > 0, also exceeds
    adapt times the median NRMSE of the last 5 simulated windows (the error
    floor that a single noisy realisation sets), so that fallbacks follow
    drift of the surrogate rather than the noise of the simulator."""

    def __init__(self, da, surrogate, every=5, nrmse_max=0.02, finetune_epochs=None, replay=1.0, adapt=0.0):
        self.da = da
        self.sur = surrogate
        self.every = int(every)
        self.nrmse_max = float(nrmse_max)
        self.adapt = float(adapt)
        if finetune_epochs is None:
            finetune_epochs = 4 if getattr(surrogate, "per_voxel", False) else 40
        self.finetune_epochs = int(finetune_epochs)
        self.replay = float(replay)

    def _advance_hp(self, hp_sigma):
        da = self.da
        da._hp_log += torch.normal(0, hp_sigma ** 0.5, size=da._hp_log.shape).float()
        da._hp = da.sigmoid_torch(da._hp_log, da._hp_low, da._hp_high)

    def run(self, observe, windows, steps=800, bold_sigma=1e-8, solo_rate=0.8, hp_sigma=0.25):
        """`observe(t)` -> observed BOLD [voxels per member].  Returns a dict with
        per-window forecasts, flags, NRMSE of simulated windows and times."""
        da = self.da
        E = da.ensemble_number
        forecasts, flags, errs, times = [], [], [], []
        force = False
        for t in range(windows):
            t0 = time.perf_counter()
            self._advance_hp(hp_sigma)
            X = window_inputs(da)
            simulate = force or (t % self.every == 0)
            pred = self.sur.predict(X)
            if simulate:
                da.push_hp()
                da.get_hidden_state(steps)
                tail = da._hidden_state[:, :, -6:].double().cpu()
                e = nrmse(pred[:, :, -1].cpu().numpy(), tail[:, :, -1].numpy())
                floor = float(np.median([x for _, x in errs[-5:]])) if (self.adapt > 0 and errs) else 0.0
                errs.append((t, e))
                force = e > max(self.nrmse_max, self.adapt * floor)
                self.sur.add(X, tail.reshape(E, -1))
                if self.finetune_epochs:
                    self.sur.finetune(self.finetune_epochs, replay=self.replay, seed=t)
            else:
                hpl = da._hp_log.reshape(E, -1)[:, da.for_hidden_state].reshape(E, pred.shape[1], -1)
                da._hidden_state = torch.cat((hpl.double().to(pred.device), pred), dim=2)
            forecasts.append(da._hidden_state.cpu().numpy())
            flags.append(simulate)
            y = observe(t)
            da.da_filter(torch.as_tensor(y).reshape(1, -1), bold_sigma=bold_sigma, solo_rate=solo_rate)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return dict(forecasts=np.stack(forecasts), simulated=np.array(flags), nrmse=errs, times=np.array(times))


def collect_windows(da, windows, steps, logit_range=10.0, rng=None, walk_sigma=None):
    """Simulator windows without analysis; returns (X [windows*E, 8V],
    Y [windows*E, 6V]) for training.  Hyper-parameter logits are either
    fresh uniform draws from [-logit_range, logit_range] each window
    (walk_sigma None), or a Gaussian random walk of step walk_sigma from the
    current logits, the way the assimilation moves them."""
    rng = np.random.default_rng(0) if rng is None else rng
    Xs, Ys = [], []
    E = da.ensemble_number
    for _ in range(windows):
        if walk_sigma is None:
            hl = rng.uniform(-logit_range, logit_range, size=tuple(da._hp_log.shape))
        else:
            hl = da._hp_log.numpy() + rng.normal(0.0, walk_sigma, size=tuple(da._hp_log.shape))
        da._hp_log = torch.from_numpy(hl).float()
        da._hp = da.sigmoid_torch(da._hp_log, da._hp_low, da._hp_high)
        X = window_inputs(da)
        da.push_hp()
        da.get_hidden_state(steps)
        Xs.append(X)
        Ys.append(da._hidden_state[:, :, -6:].double().cpu().reshape(E, -1))
    return torch.cat(Xs), torch.cat(Ys)
