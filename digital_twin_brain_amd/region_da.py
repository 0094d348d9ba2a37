"""Surrogate-assisted data assimilation on an ensemble of whole-brain members
(no counterpart in the reference: the surrogate, the judge and the fallback
policy are new in this library; the assimilation loop they accelerate is
data_assimilation/DataAssimilation.py:402-453).

Members.  Every member is a replica of one R-region network (lognormal region
sizes, intra- and inter-region inputs through one shared connectome, delays
up to 40 steps, external Poisson drive), packed with a hidden truth replica
into one block (networks.build_region_ensemble).  Its hyper-parameters are the
4 receptor conductances g_ui (property columns 10-13) of each region, shared
by the region's E and I populations: the hidden state per voxel (= region) is
[4 hp logits, act, s, q, v, f_in, bold].

Surrogate.  The window map from the 4 hp logits of every region, the BOLD
state (s, q, v, f_in) at the window start and the previous window's rate per
region to the window's (act, s, q, v, f_in, bold) per region, learned by a
3 x 512 bf16 MLP on the simulator's own windows.  The default form
(WindowSurrogate per_voxel=True) applies one MLP to every region: its input is
the region's own hp logits and BOLD state and the member's mean of them, so
every window gives E x R training rows.  The previous window's rate is not
an input: the hidden state's act is the last step's rate (mostly noise), and
after a surrogate window it would be the surrogate's own smooth prediction,
unlike anything in training.  On 50,176 training windows (256 members) it
predicts held-out windows' BOLD at about 1% NRMSE; the per-member form (all
810 features of a member in, 540 out, per_voxel=False) reached about 3%,
with or without a one-hot of the member.

Fallback (member-subset judge).  A forced full-ensemble simulation every
`every`-th window caps the gain below `every` (the window cost is then at
least T_sim / every).  Here the periodic check simulates only the first
`subset` members, on a separate block holding copies of exactly those
members (the same synapses and properties, another noise stream:
SubsetJudge): it pushes their current hyper-parameters, lets the network
relax under them for `pre_ms`, sets their window-start BOLD state and runs
the window.  Its cost is T_sim * subset / E.  The subset's simulated windows
replace the surrogate's forecasts for those members, are added to the
training set and fine-tune the surrogate.  When the subset's NRMSE
(surrogate against simulation, over the subset's BOLD) exceeds `nrmse_max`,
the next window simulates the full ensemble, and so on until a full window
passes.

The simulator is stochastic: two simulations of the same members from the
same state and hyper-parameters differ.  run_region_surrogate_da measures
that floor (the judge's simulation against the full simulation of the same
members in one window) and, with nrmse_max=None, sets the threshold to
max(0.02, 1.25 x floor): the surrogate counts as wrong when it departs from
the simulation by more than another simulation would.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import networks
from .data_assimilation import DataAssimilation
from .surrogate_da import NF as NF_DA, WindowSurrogate, nrmse, window_inputs

HP_COLUMNS = (10, 11, 12, 13)  # g_AMPA, g_NMDA, g_GABAa, g_GABAb of every region


def make_region_da(E, R, member_n, dt=0.1, hp_lo=0.5, hp_hi=2.0, seed=7, truth=True, tune=True, **kw):
    """Region ensemble (+ truth replica) and its DataAssimilation; the initial
    hyper-parameters are the reference's logit grid, pushed; the truth's
    factors are drawn once and pushed.  Returns (nb, da, truth_hp, (low, high))."""
    rng = np.random.default_rng(seed)
    nb, props, spec, cdf = networks.build_region_ensemble(E=E, R=R, member_n=member_n, truth=truth, dt=dt, **kw)
    sps = int(round(1.0 / dt))  # BOLD once per ms
    da = DataAssimilation(network=nb, column=False, ensemble=E, dt=dt, bold_steps=sps, da_populations=E * 2 * R,
                          da_voxels=E * R, write_path="/tmp")
    H = len(HP_COLUMNS)
    low = torch.full((R, H), hp_lo)
    high = torch.full((R, H), hp_hi)
    hp = da.hp_random_initialize(low, high, num_da_populations_pblk=R, rng=rng)
    da.hp_index2hidden_state(hp_rows_per_voxel=1)
    da.da_property_initialize(list(HP_COLUMNS), None, hp, pops_per_hp=2)
    truth_hp = None
    if truth:
        truth_hp = rng.uniform(hp_lo, hp_hi, size=(R, H)).astype(np.float32)
        pops = np.arange(E * 2 * R, (E + 1) * 2 * R)
        nb.mul_props_by_pop(np.repeat(pops, H).astype(np.uint32), np.tile(HP_COLUMNS, 2 * R).astype(np.uint32),
                            np.repeat(truth_hp, 2, axis=0).reshape(-1))
    if tune:
        from .autotune import tune_delivery

        nb.run(10 * sps, noise_mode=1, graph=False)  # leave the initial transient before measuring
        tune_delivery(nb, dict(noise_mode=1), steps=2 * sps)
    return nb, da, truth_hp, (low, high)


class RateSurrogate:
    """Window forecast from a learned window-mean BOLD input.  An MLP maps a
    region's 4 hyper-parameter logits, its member's mean logits and a one-hot
    of the region to log of the window's mean BOLD input u (spikes per BOLD
    step over the region's neurons); the hemodynamics are then integrated
    from the window-start state (s, q, v, f_in) with that constant input,
    BOLD step by BOLD step (dtb_bold_run_f64, the Balloon-Windkessel model of
    the simulation), and act is u per network step.  The hemodynamic state
    so stays on the model's own trajectories across consecutive surrogate
    windows; only the rate is learned.  Same interface as WindowSurrogate
    (add / fit / finetune / predict), the training target being u."""

    def __init__(self, voxels, bold_steps, window_bold_steps, hidden=(256, 256), max_batch=2048, lr=1e-3, seed=0,
                 device=0):
        from .simulation import bold_params
        from .surrogate import MLPSurrogate

        self.V = int(voxels)
        self.sps = int(bold_steps)           # network steps per BOLD step
        self.nbs = int(window_bold_steps)    # BOLD steps per window
        p = bold_params
        self.par = torch.tensor([p["epsilon"], p["tao_s"], p["tao_f"], p["tao_0"], p["alpha"], p["E_0"], p["V_0"],
                                 1e-3], dtype=torch.float64, device=torch.device("cuda", device))
        self.mlp = MLPSurrogate(8 + self.V, 1, hidden=hidden, max_batch=max_batch, lr=lr, seed=seed, device=device)
        self.X = []
        self.Y = []

    def _rows(self, X):
        x = torch.as_tensor(X, dtype=torch.float64).cpu().reshape(-1, self.V, NF_DA)
        hp = x[..., :4]
        m = hp.mean(dim=1, keepdim=True).expand(-1, self.V, -1)
        ids = torch.eye(self.V, dtype=torch.float64).expand(x.shape[0], -1, -1)
        return torch.cat((hp, m, ids), dim=2).reshape(x.shape[0] * self.V, -1)

    def add(self, X, Y=None, idx=None, u=None):
        """X [E, 9V] window inputs; u [E, V] the simulated window-mean input."""
        u = torch.as_tensor(u, dtype=torch.float64).cpu().reshape(-1, 1)
        self.X.append(self._rows(X))
        self.Y.append(torch.log(u.clamp_min(1e-6)))

    def fit(self, epochs=None, batch=None, renormalize=True):
        X = torch.cat(self.X)
        Y = torch.cat(self.Y)
        if renormalize:
            self.mlp.set_normalization(X, Y)
        return self.mlp.fit(X, Y, epochs=epochs or 20, batch=batch)

    finetune = WindowSurrogate.finetune

    def rate(self, X):
        """[E, V] predicted window-mean input."""
        E = torch.as_tensor(X).shape[0]
        return torch.exp(self.mlp.predict(self._rows(X)).double()).reshape(E, self.V)

    def predict(self, X, idx=None):
        from . import _native as nat

        X = torch.as_tensor(X, dtype=torch.float64)
        E = X.shape[0]
        dev = self.par.device
        u = self.rate(X).to(dev).clamp_min(1e-5).contiguous().reshape(-1)
        x = X.reshape(E, self.V, NF_DA)
        st = [x[..., c].reshape(-1).to(dev).contiguous() for c in (4, 5, 6, 7)]  # s, q, v, f_in
        out = torch.zeros(E * self.V, dtype=torch.float64, device=dev)
        nat.check(nat.lib().dtb_bold_run_f64(E * self.V, self.nbs, self.par.data_ptr(), st[0].data_ptr(),
                                              st[1].data_ptr(), st[2].data_ptr(), st[3].data_ptr(), u.data_ptr(),
                                              out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream))
        act = u / self.sps
        return torch.stack((act, st[0], st[1], st[2], st[3], out), dim=1).reshape(E, self.V, 6)


class SubsetJudge:
    """Copies of the first `subset` members on a block of their own (the
    generator draws a member's synapses from its own neuron indices, so the
    same seeds give the same members), with another noise stream; simulates
    their windows from given hyper-parameters and window-start BOLD state."""

    def __init__(self, subset, R, member_n, dt=0.1, hp_lo=0.5, hp_hi=2.0, seed=8, pre_ms=50.0, noise_seed=103,
                 **kw):
        self.S = int(subset)
        self.nb, self.da, _, _ = make_region_da(self.S, R, member_n, dt=dt, hp_lo=hp_lo, hp_hi=hp_hi, seed=seed,
                                                truth=False, noise_seed=noise_seed, **kw)
        self.pre_steps = int(round(pre_ms / dt))
        self.pre_steps -= self.pre_steps % self.da.bold_steps

    def simulate(self, hp_log, bold_start, steps, with_input=False):
        """hp_log [S, R*4] logits; bold_start [S, R, 4] (s, q, v, f_in) ->
        the window's hidden-state tail [S, R, 6] (float64, CPU), and with
        with_input also its mean BOLD input per region [S, R]."""
        da = self.da
        da._hp_log = torch.as_tensor(hp_log, dtype=torch.float32).reshape(self.S, -1).clone()
        da._hp = da.sigmoid_torch(da._hp_log, da._hp_low, da._hp_high)
        da.push_hp()
        if self.pre_steps:
            da.evolve(self.pre_steps, bold_detail=False)  # the network settles under the new conductances
        da.bold.state_update(torch.as_tensor(bold_start, dtype=torch.float64))
        da.get_hidden_state(steps, window_input=with_input)
        tail = da._hidden_state[:, :, -6:].double().cpu()
        if with_input:
            return tail, da.last_window_input[: tail.shape[0] * tail.shape[1]].reshape(tail.shape[:2]).cpu()
        return tail


class SubsetSurrogateDA:
    """The window loop with the surrogate forecasting every member, a
    member subset simulated every `every`-th window as the judge,
    and full-ensemble simulation after a judged (or full) window whose NRMSE
    exceeds nrmse_max.  Every simulated window fine-tunes the surrogate on
    its rows plus an equal replay of earlier rows."""

    def __init__(self, da, surrogate, judge, every=5, nrmse_max=0.02, finetune_epochs=4, replay=1.0,
                 residual=None, seed=0, rejudge=0, nrmse_scale="members", refull=False):
        self.da = da
        # refull=False: a full window's NRMSE is recorded, but only judged
        # windows send the next window to the simulator.  A full window starts
        # from the surrogate-driven state of the windows before it (that error
        # is what its NRMSE measures), and it hands the next window a
        # simulated state, from which the surrogate starts out as accurate as
        # after any simulated window.  refull=True: a full window over the
        # threshold is followed by another full window.
        self.refull = bool(refull)
        # a surrogate trained on the window-mean input needs it from every simulated window
        self.with_input = isinstance(surrogate, RateSurrogate)
        # a judged window's NRMSE divides the RMSE over the judged members by the
        # range of the window's BOLD over the whole ensemble ("ensemble": the
        # judged members' simulated values, the others' forecasts), the scale a
        # full window's NRMSE has; "members": by the judged members' own range
        self.nrmse_scale = nrmse_scale
        # a judged window over the threshold is followed by up to `rejudge`
        # further judged windows (the surrogate fine-tuned on each) before the
        # full ensemble is simulated
        self.rejudge = int(rejudge)
        # residual [V, 6]: the simulator's spread around the surrogate's mean; surrogate
        # forecasts that enter the analysis carry draws of it, so that the ensemble's
        # forecast spread, and with it the Kalman gain, is that of simulated windows
        self.residual = None if residual is None else torch.as_tensor(residual, dtype=torch.float64)
        self.gen = torch.Generator().manual_seed(int(seed))
        self.sur = surrogate
        self.judge = judge
        self.every = int(every)
        self.nrmse_max = float(nrmse_max)
        self.finetune_epochs = int(finetune_epochs)
        self.replay = float(replay)

    def run(self, observe, windows, steps, bold_sigma=1e-8, solo_rate=0.8, hp_sigma=0.25):
        da = self.da
        E = da.ensemble_number
        S = self.judge.S
        forecasts, kinds, errs, times = [], [], [], []
        phases = []  # per window: seconds in (forecast, simulation, fine-tuning, analysis)
        full_next = judge_next = False
        fails = 0  # consecutive judged windows over the threshold
        for t in range(windows):
            t0 = time.perf_counter()
            da._hp_log += torch.normal(0, hp_sigma ** 0.5, size=da._hp_log.shape).float()
            da._hp = da.sigmoid_torch(da._hp_log, da._hp_low, da._hp_high)
            X = window_inputs(da)
            pred = self.sur.predict(X)  # [E, V, 6] on the device: the surrogate's mean
            hpl = da._hp_log.reshape(E, -1)[:, da.for_hidden_state].reshape(E, pred.shape[1], -1)
            fc = pred
            if self.residual is not None:
                draw = torch.randn(pred.shape, generator=self.gen, dtype=torch.float64) * self.residual
                fc = pred + draw.to(pred.device)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if full_next:
                kind = "full"
                da.push_hp()
                if kinds and kinds[-1] != "full" and self.judge.pre_steps:
                    # the ensemble last ran windows ago, under other conductances: it settles
                    # first, then starts from the window's BOLD state, like the judge
                    da.evolve(self.judge.pre_steps, bold_detail=False)
                    _set_member_bold(da, X.reshape(E, -1, 9)[:, :, 4:8])
                da.get_hidden_state(steps, window_input=self.with_input)
                tail = da._hidden_state[:, :, -6:].double().cpu()
                e = nrmse(pred[:, :, -1].cpu().numpy(), tail[:, :, -1].numpy())
                u = da.last_window_input[: E * tail.shape[1]].reshape(E, -1).cpu() if self.with_input else None
                self.sur.add(X, tail.reshape(E, -1), u=u)
            elif t % self.every == 0 or judge_next:
                kind = "judge"
                J = np.arange(S)
                bold0 = X.reshape(E, -1, 9)[J][:, :, 4:8]
                u = None
                if self.with_input:
                    tail, u = self.judge.simulate(da._hp_log.reshape(E, -1)[J], bold0, steps, with_input=True)
                else:
                    tail = self.judge.simulate(da._hp_log.reshape(E, -1)[J], bold0, steps)
                self.sur.add(X[J], tail.reshape(S, -1), idx=J, u=u)
                full = torch.cat((hpl.double().to(pred.device), fc), dim=2)
                full[torch.from_numpy(J).to(pred.device), :, -6:] = tail.to(pred.device)
                da._hidden_state = full
                scale = None
                if self.nrmse_scale == "ensemble":
                    wb = full[:, :, -1]
                    scale = float((wb.max() - wb.min()).item())
                e = nrmse(pred[J][:, :, -1].cpu().numpy(), tail[:, :, -1].numpy(), scale)
            else:
                kind = "surrogate"
                e = None
                da._hidden_state = torch.cat((hpl.double().to(pred.device), fc), dim=2)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if e is not None:
                errs.append((t, kind, e))
                bad = e > self.nrmse_max
                fails = fails + 1 if (bad and kind == "judge") else 0
                judge_next = bad and kind == "judge" and fails <= self.rejudge
                full_next = bad and not judge_next and (kind == "judge" or self.refull)
                if self.finetune_epochs:
                    # a full window adds E x R rows: a sample of them keeps the update cheap
                    self.sur.finetune(self.finetune_epochs, replay=self.replay, seed=t, max_rows=4 * 1024)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            forecasts.append(da._hidden_state.cpu().numpy())
            kinds.append(kind)
            y = observe(t)
            da.da_filter(torch.as_tensor(y).reshape(1, -1), bold_sigma=bold_sigma, solo_rate=solo_rate)
            torch.cuda.synchronize()
            t4 = time.perf_counter()
            times.append(t4 - t0)
            phases.append((t1 - t0, t2 - t1, t3 - t2, t4 - t3))
        return dict(forecasts=np.stack(forecasts), kinds=kinds, nrmse=errs, times=np.array(times),
                    phases=np.array(phases))


def collect_da_windows(da, windows, steps, rng, low, high, R, segment=8, hp_sigma=0.25, bold_sigma=1e-8,
                       solo_rate=0.8, seed=0, log=None, with_input=False):
    """Training windows drawn the way the assimilation moves: segments of
    `segment` windows that start from the reference's logit grid, take the
    assimilation's random-walk step and then an EnKF analysis each window,
    against a pseudo-observation: the simulated BOLD of one member picked at
    random per segment, whose hyper-parameters are drawn uniformly from the
    prior range and held fixed (a stand-in truth, never the real one), so
    that the ensemble's hyper-parameters and BOLD states travel as they do
    when assimilating an observation from outside the ensemble.
    Returns (X [windows*E, 9R], Y [windows*E, 6R]), and with with_input also
    U [windows*E, R], each window's mean BOLD input per region."""
    E = da.ensemble_number
    Xs, Ys, Us = [], [], []
    for w0 in range(0, windows, segment):
        da.hp_random_initialize(low, high, num_da_populations_pblk=R, rng=np.random.default_rng(seed + w0))
        target = int(rng.integers(E))
        lo, hi = da._hp_low.numpy(), da._hp_high.numpy()
        u = rng.uniform(0.02, 0.98, size=lo.shape)
        fixed = torch.from_numpy(10.0 * np.log(u / (1.0 - u))).float()  # logit of lo + (hi - lo) u
        for _ in range(min(segment, windows - w0)):
            da._hp_log += torch.from_numpy(rng.normal(0.0, hp_sigma ** 0.5, size=tuple(da._hp_log.shape))).float()
            da._hp_log[target] = fixed
            da._hp = da.sigmoid_torch(da._hp_log, da._hp_low, da._hp_high)
            X = window_inputs(da)
            da.push_hp()
            da.get_hidden_state(steps, window_input=with_input)
            Xs.append(X)
            Ys.append(da._hidden_state[:, :, -6:].double().cpu().reshape(E, -1))
            if with_input:
                Us.append(da.last_window_input[: E * R].reshape(E, R).cpu())
            y = da._hidden_state[target, :, -1].reshape(1, -1)
            da.da_filter(y, bold_sigma=bold_sigma, solo_rate=solo_rate)
        if log:
            log(f"region DA: {len(Xs)} of {windows} training windows")
    if with_input:
        return torch.cat(Xs), torch.cat(Ys), torch.cat(Us)
    return torch.cat(Xs), torch.cat(Ys)


def _set_member_bold(da, state):
    """BOLD state (s, q, v, f_in) [E, R, 4] of the members' voxels; voxels past
    the members (the truth) keep theirs."""
    nb = da.block_model.nb
    full = np.stack([nb.bold_state(w) for w in ("s", "q", "v", "f_in")], axis=1)
    full[: da._da_voxels] = np.asarray(state, np.float64).reshape(-1, 4)
    da.bold.state_update(torch.from_numpy(full).reshape(1, -1, 4))


def _snapshot(da):
    return dict(block=da.block_model.state_dict(), hp_log=da._hp_log.clone(), hp=da._hp.clone(),
                hidden=None if da._hidden_state is None else da._hidden_state.clone(),
                rng_cpu=torch.get_rng_state(), rng_dev=torch.cuda.get_rng_state(da.block_model._dev))


def _restore(da, s):
    da.block_model.load_state_dict(s["block"])
    da._hp_log = s["hp_log"].clone()
    da._hp = s["hp"].clone()
    da._hidden_state = None if s["hidden"] is None else s["hidden"].clone()
    torch.set_rng_state(s["rng_cpu"])
    torch.cuda.set_rng_state(s["rng_dev"], da.block_model._dev)
    da.timings = []


def run_region_surrogate_da(E=256, R=90, member_n=10_000, train_windows=196, windows=20, window_ms=800.0, dt=0.1,
                            subset=16, every=5, nrmse_max=None, hidden=(512, 512, 512), epochs=None,
                            finetune_epochs=None, seed=7, pre_ms=50.0, per_voxel=True, sample_residual=False, log=None,
                            tune=True, rejudge=0, nrmse_scale="members", refull=False, calib_windows=3,
                            surrogate="window", rate_hidden=(256, 256), **kw):
    """Simulator-only DA against surrogate-assisted DA (subset judge) on E
    whole-brain members, from the same state and hyper-parameters, against
    the same truth observations.  Returns a dict of timings, gain, the
    fallback pattern and NRMSE values, and the run records."""
    say = log or (lambda *a: None)
    t0 = time.perf_counter()
    nb, da, truth_hp, (low, high) = make_region_da(E, R, member_n, dt=dt, seed=seed, tune=tune, **kw)
    judge = SubsetJudge(subset, R, member_n, dt=dt, seed=seed + 1, pre_ms=pre_ms, **kw)
    t_build = time.perf_counter() - t0
    steps = int(round(window_ms / dt))
    truth_vox = slice(E * R, (E + 1) * R)
    da.evolve(steps, bold_detail=True)
    say(f"region DA: {E} members + truth, {nb.n} neurons, {nb.nnz} synapses, judge {subset} members, built and "
        f"warmed in {time.perf_counter() - t0:.1f} s")
    # training windows: assimilation-like segments against pseudo-observations
    t0 = time.perf_counter()
    rate = surrogate == "rate"
    got = collect_da_windows(da, train_windows, steps, np.random.default_rng(seed + 3), low, high, R,
                             seed=seed + 100, log=log, with_input=rate)
    X, Y = got[0], got[1]
    t_collect = time.perf_counter() - t0
    say(f"region DA: {train_windows} training windows ({X.shape[0]} samples) in {t_collect:.1f} s")
    if rate:
        if sample_residual:
            raise ValueError("residual draws are for the window surrogate")
        ws = RateSurrogate(R, da.bold_steps, steps // da.bold_steps, hidden=rate_hidden, seed=seed)
        ws.add(X, u=got[2])
    else:
        ws = WindowSurrogate(R, hidden=hidden, seed=seed, per_voxel=per_voxel, max_batch=2048 if per_voxel else 256,
                             use_act=False)
        ws.add(X, Y)
    torch.cuda.synchronize()
    if epochs is None:
        epochs = 5 if per_voxel else 60
    if finetune_epochs is None:
        finetune_epochs = 4 if per_voxel else 20
    t0 = time.perf_counter()
    hist = ws.fit(epochs=epochs)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    say(f"region DA: surrogate fitted in {t_fit:.1f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    nres = min(X.shape[0], 16 * E)  # the spread of the last training windows around the fit
    residual = ws.residual_std(X[-nres:], Y[-nres:]) if sample_residual else None
    # the simulator's own floor: the judge and the full ensemble simulate the first members' window
    # from the same state and hyper-parameters (two noise streams), over `calib_windows`
    # consecutive windows; the threshold uses the median
    J = np.arange(subset)
    cal = []
    for _ in range(max(1, int(calib_windows))):
        Xc = window_inputs(da)
        tail_j = judge.simulate(da._hp_log.reshape(E, -1)[J], Xc.reshape(E, -1, 9)[J][:, :, 4:8], steps)
        da.push_hp()
        da.get_hidden_state(steps)
        tail_m = da._hidden_state[:, :, -6:].double().cpu()
        pred_c = ws.predict(Xc)[:, :, -1].cpu().numpy()
        # judged windows' scale: the range over the whole ensemble's window
        sc = float(tail_m[:, :, -1].max() - tail_m[:, :, -1].min()) if nrmse_scale == "ensemble" else None
        cal.append(dict(sim_vs_sim=nrmse(tail_j[:, :, -1].numpy(), tail_m[J][:, :, -1].numpy(), sc),
                        sur_vs_full=nrmse(pred_c[J], tail_m[J][:, :, -1].numpy(), sc),
                        sur_vs_judge=nrmse(pred_c[J], tail_j[:, :, -1].numpy(), sc),
                        sur_vs_full_all_members=nrmse(pred_c, tail_m[:, :, -1].numpy())))
    calib = {k: float(np.median([c[k] for c in cal])) for k in cal[0]}
    calib["sim_vs_sim_windows"] = [round(c["sim_vs_sim"], 5) for c in cal]
    floor = calib["sim_vs_sim"]
    if nrmse_max is None:
        nrmse_max = max(0.02, 1.25 * floor)
    say(f"region DA: NRMSE floor (simulation against simulation) {floor:.4f}, surrogate {calib}; threshold "
        f"{nrmse_max:.4f}")
    # both runs start from this state and these hyper-parameters
    hp = da.hp_random_initialize(low, high, num_da_populations_pblk=R, rng=np.random.default_rng(seed + 1))
    da._push(hp)
    da._hidden_state = None
    snap = _snapshot(da)
    obs = []

    def observe(t):
        y = nb.bold_state("bold")[truth_vox].copy()
        obs.append(y)
        return torch.from_numpy(y)

    da.timings = []
    w_sim, _, t_sim = da.da_rest_run(observe, "/tmp", observation_times=windows, steps=steps, save=False)
    obs = np.stack(obs)
    say(f"region DA: {windows} simulator-only windows, {t_sim.mean():.3f} s each")
    _restore(da, snap)
    del snap
    sda = SubsetSurrogateDA(da, ws, judge, every=every, nrmse_max=nrmse_max, finetune_epochs=finetune_epochs,
                            residual=residual, seed=seed, rejudge=rejudge, nrmse_scale=nrmse_scale,
                            refull=refull)
    res = sda.run(lambda t: torch.from_numpy(obs[t]), windows, steps=steps)
    kinds = res["kinds"]
    say(f"region DA: {windows} surrogate-assisted windows, {res['times'].mean():.3f} s each, "
        f"{kinds.count('judge')} judged, {kinds.count('full')} full")
    rmse = lambda w: np.sqrt(((w[:, :, :, -1].mean(1) - obs) ** 2).mean(1))
    errs = [e for _, _, e in res["nrmse"]]
    out = dict(members=E, regions=R, member_neurons=member_n, windows=windows, steps_per_window=steps,
               nrmse_threshold=nrmse_max,
               calibration={k: (round(v, 5) if isinstance(v, float) else v) for k, v in calib.items()},
               neurons=int(nb.n), synapses=int(nb.nnz), judge_members=subset, judge_every=every, rejudge=rejudge,
               nrmse_scale=nrmse_scale, refull=refull,
               build_s=t_build, train_windows=train_windows, train_samples=int(X.shape[0]), collect_s=t_collect,
               fit_s=t_fit, fit_epochs=epochs, fit_loss_first=hist[0], fit_loss_last=hist[-1],
               surrogate=(f"per-region rate MLP {8 + R} -> " + "x".join(str(h) for h in rate_hidden) +
                          " -> log mean BOLD input, hemodynamics integrated" if rate else
                          (f"per-region MLP {2 * 9} -> " if per_voxel else f"per-member MLP {R * 9} -> ") +
                          "x".join(str(h) for h in hidden) + (" -> 6" if per_voxel else f" -> {R * 6}")),
               sim_window_s_mean=float(t_sim.mean()), sur_window_s_mean=float(res["times"].mean()),
               gain=float(t_sim.sum() / res["times"].sum()), kinds=kinds,
               window_s_by_kind={k: float(np.mean([x for x, kk in zip(res["times"], kinds) if kk == k]))
                                 for k in sorted(set(kinds))},
               window_phase_s_by_kind={k: dict(zip(("forecast", "simulation", "finetune", "analysis"),
                                                   np.round(res["phases"][[kk == k for kk in kinds]].mean(0), 4)
                                                   .tolist()))
                                       for k in sorted(set(kinds))},
               nrmse=[(t, k, round(e, 5)) for t, k, e in res["nrmse"]],
               nrmse_median=float(np.median(errs)) if errs else None, nrmse_max=float(np.max(errs)) if errs else None,
               bold_rmse_sim_last=float(rmse(w_sim)[-1]), bold_rmse_sur_last=float(rmse(res["forecasts"])[-1]),
               bold_rmse_sim_mean=float(rmse(w_sim).mean()), bold_rmse_sur_mean=float(rmse(res["forecasts"]).mean()))
    say(f"region DA: {out}")
    return out, dict(w_sim=w_sim, sur=res, obs=obs, hist=hist, truth_hp=truth_hp, da=da, judge=judge)
