"""In-process `DataAssimilation` (data_assimilation/DataAssimilation.py:63-453).

Same methods and hidden-state layout as the reference class:
- hyper-parameters follow a logit random walk, `hp = lo + (hi-lo) sigmoid(hp_log/10)`;
- property pushes are multiplicative per sub-population;
- the hidden state of a window is `[hp_log..., act, s, q, v, f_in, bold]` per voxel;
- the diffusion EnKF runs in FP64 on the GPU;
- the BOLD state is written back after the analysis.

Members are the first `ensemble` equal slices of the block's sub-populations
and voxels; a block may hold more sub-populations after them (for example a
hidden "truth" network producing the observed BOLD), which the assimilation
ignores.  Plotting is left to the caller.
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch

from .enkf import diffusion_enkf
from .simulation import simulation


class DataAssimilation(simulation):
    def __init__(self, block_path=None, ip=None, dt=1.0, route_path=None, column=True, ensemble=100, **kwargs):
        super().__init__(ip, block_path, dt, route_path, column, **kwargs)
        self._ensemble_number = int(ensemble)
        self._da_populations = int(kwargs.get("da_populations", self.num_populations))
        self._da_voxels = int(kwargs.get("da_voxels", self.num_voxels))
        assert self._da_populations % self._ensemble_number == 0
        assert self._da_voxels % self._ensemble_number == 0
        self._hidden_state = None
        self._property_index = None
        self._index_da_voxel_pblk = None
        self._hp_index_updating = None
        self._hp_log = None
        self._hp = None
        self._hp_low = None
        self._hp_high = None
        self._hp_num = None
        self.for_hidden_state = None
        self.from_hidden_state = None
        self.timings = []

    # data_assimilation/DataAssimilation.py:110-140
    @staticmethod
    def log_torch(val, lower, upper, scale=10):
        shp = val.shape
        val = val.reshape(-1, lower.shape[-1])
        return (scale * (torch.log(val - lower) - torch.log(upper - val))).reshape(shp)

    @staticmethod
    def sigmoid_torch(val, lower, upper, scale=10):
        shp = val.shape
        val = val.reshape(-1, lower.shape[-1])
        return (lower + (upper - lower) * torch.sigmoid(val / scale)).reshape(shp)

    @property
    def ensemble_number(self):
        return self._ensemble_number

    @property
    def _num_populations_pblk(self):
        return self._da_populations // self._ensemble_number

    @property
    def _num_voxel_pblk(self):
        return self._da_voxels // self._ensemble_number

    def hp_random_initialize(self, gui_low_ppopu, gui_high_ppopu, num_da_populations_pblk=None, gui_pblk=False,
                             rng=None):
        """data_assimilation/DataAssimilation.py:204-252: a logit-spaced grid over
        (-1/3, 1/3) per hyper-parameter, members shuffled independently per column."""
        rng = np.random.default_rng() if rng is None else rng
        P = self._num_populations_pblk if num_da_populations_pblk is None else num_da_populations_pblk
        if gui_pblk:
            gui_low_ppopu = gui_low_ppopu.repeat(P, 1)
            gui_high_ppopu = gui_high_ppopu.repeat(P, 1)
        assert gui_low_ppopu.shape[0] == P
        E = self._ensemble_number
        self._hp_num = gui_low_ppopu.shape[1]
        self._hp_low = gui_low_ppopu.reshape(-1).float()
        self._hp_high = gui_high_ppopu.reshape(-1).float()
        temp = torch.linspace(-1 / 3, 1 / 3, E)
        hp_log = (10 * (torch.log(temp + 1) - torch.log(1 - temp))).repeat_interleave(len(self._hp_low))
        hp_log = hp_log.float().reshape(E, -1, self._hp_num)
        for i in range(self._hp_num):
            idx = torch.from_numpy(rng.permutation(E))
            hp_log[:, :, i] = hp_log[idx, :, i]
        self._hp_log = hp_log.reshape(E, -1)
        self._hp = self.sigmoid_torch(self._hp_log, self._hp_low, self._hp_high).float()
        return self._hp.reshape(-1)

    def hp_index2hidden_state(self, cortical_or_not=None, index_da_voxel_pblk=None, hp_rows_per_voxel=2):
        """data_assimilation/DataAssimilation.py:254-293 (voxel version: 2 populations
        per voxel; micro-columns: 8 cortical + 2 sub-cortical layouts).  With
        hp_rows_per_voxel=1 one row of hyper-parameters serves both
        populations of a voxel (see da_property_initialize's pops_per_hp)."""
        Vb = self._num_voxel_pblk
        self._index_da_voxel_pblk = torch.arange(Vb) if index_da_voxel_pblk is None else \
            torch.tensor(index_da_voxel_pblk).reshape(-1)
        cort = np.zeros(Vb, bool) if cortical_or_not is None else np.asarray(cortical_or_not).reshape(-1) == 1
        cort = cort[self._index_da_voxel_pblk.numpy()]
        nv = cort.shape[0]
        H = self._hp_num
        if cort.sum() == 0:
            self.for_hidden_state = torch.arange(nv * hp_rows_per_voxel * H)
            self.from_hidden_state = torch.arange(nv * hp_rows_per_voxel * H)
        else:
            fhs = 10 * torch.arange(nv).reshape(-1, 1) + torch.tensor([6, 7]).repeat(4)
            fhs[torch.from_numpy(cort)] += torch.tensor([-4, -4, -2, -2, 0, 0, 2, 2])
            self.for_hidden_state = fhs.reshape(-1).repeat_interleave(H)
            vox = (8 * torch.arange(nv).reshape(-1, 1) + torch.tensor([0, 1])).reshape(-1)
            cor = (8 * torch.arange(nv)[torch.from_numpy(cort)].reshape(-1, 1) + torch.tensor([2, 3, 4, 5, 6, 7]))
            frm, _ = torch.cat((vox, cor.reshape(-1))).reshape(-1, 1).sort(dim=0)
            self.from_hidden_state = (H * frm + torch.arange(H)).reshape(-1)

    def da_property_initialize(self, property_index, alpha, gui, index_da_population=None, pops_per_hp=1):
        """data_assimilation/DataAssimilation.py:295-325.  pops_per_hp > 1: one
        row of hyper-parameters (gui) drives that many consecutive populations
        (for example a region's E and I populations sharing its 4 g_ui)."""
        pops = self.population_id.cpu()[: self._da_populations] if index_da_population is None \
            else torch.as_tensor(index_da_population)
        if alpha is not None:  # alpha=None keeps the existing values and only scales them
            for p in property_index:
                self.gamma_initialize(p, population_id=pops, alpha=alpha, beta=alpha)
        self._property_index = torch.tensor(property_index, dtype=torch.int64).reshape(-1)
        self._hp_index_updating = torch.stack(torch.meshgrid(pops, self._property_index, indexing="ij"),
                                              dim=-1).reshape(-1, 2).to(torch.int64)
        H = self._property_index.numel()
        if pops_per_hp > 1:
            row = torch.arange(pops.numel()) // int(pops_per_hp)
            self._hp_push_gather = (row.reshape(-1, 1) * H + torch.arange(H)).reshape(-1)
        else:
            self._hp_push_gather = None
        self._push(gui)

    def _push(self, hp):
        vals = hp.float().reshape(-1)
        if getattr(self, "_hp_push_gather", None) is not None:
            vals = vals[self._hp_push_gather]
        self.mul_property_by_subblk(self._hp_index_updating, vals)

    def push_hp(self):
        """Push the current hyper-parameters into the network (relative to the
        original properties, cuda/src/stage/property.cpp:145-284)."""
        self._push(self._hp)

    def get_hidden_state(self, steps=800, show_info=False, window_input=False):
        """data_assimilation/DataAssimilation.py:327-351.  window_input: also keep
        the window's mean BOLD input per voxel (`last_window_input`)."""
        out = torch.stack(self.evolve(step=steps, bold_detail=True, window_input=window_input), dim=0).T  # [voxels, 6]
        E = self._ensemble_number
        out = out[: self._da_voxels].reshape(E, self._num_voxel_pblk, 6)[:, self._index_da_voxel_pblk]
        hp = self._hp_log.reshape(E, -1)[:, self.for_hidden_state].reshape(out.shape[0], out.shape[1], -1)
        self._hidden_state = torch.cat((hp.to(out), out), dim=2)
        self._last_window_out = out

    def da_evolve(self, steps=800, hp_sigma=0.25, generator=None):
        """data_assimilation/DataAssimilation.py:353-370."""
        self._hp_log += torch.normal(0, hp_sigma ** 0.5, size=self._hp_log.shape, generator=generator).float()
        self._hp = self.sigmoid_torch(self._hp_log, self._hp_low, self._hp_high)
        self.push_hp()
        self.get_hidden_state(steps)

    def da_filter(self, bold_real_t, bold_sigma=1e-8, solo_rate=0.8, debug=False, bound=None, xi=None):
        """data_assimilation/DataAssimilation.py:372-400."""
        if debug:
            self._hidden_state, w_debug = diffusion_enkf(self._hidden_state, bold_sigma, bold_real_t, solo_rate, True,
                                                         xi=xi)
            return w_debug
        self._hidden_state = diffusion_enkf(self._hidden_state, bold_sigma, bold_real_t, solo_rate, xi=xi)
        E = self._ensemble_number
        self._hp_log = self._hidden_state[:, :, :-6].reshape(E, -1)[:, self.from_hidden_state].float().cpu()
        if bound is not None:
            self._hp_log = torch.clamp(self._hp_log, -bound, bound)
        state = self._hidden_state[:, :, -5:-1]
        if self._da_voxels < self.num_voxels:
            # voxels outside the ensemble keep their own BOLD state
            full = torch.stack([self.bold.s, self.bold.q, self.bold.v, self.bold.f_in], dim=1)
            full[: self._da_voxels] = state.reshape(-1, 4).to(full)
            state = full.reshape(1, -1, 4)
        self.bold.state_update(state)

    # ---------------------------------------------------------- checkpoint
    def save_checkpoint(self, path, **extra):
        """Resumable state between windows: the block (state snapshot, current
        properties, relative-push factors), the hyper-parameters, the last
        hidden state and the torch RNG streams that the random walk (CPU) and
        the EnKF draw (device) use.  `extra` arrays are stored alongside.  The
        file is replaced atomically."""
        d = {"block_" + k: v for k, v in self.block_model.state_dict().items()}
        d.update(hp_log=self._hp_log.numpy(), hp=self._hp.detach().cpu().numpy(),
                 hidden_state=self._hidden_state.cpu().numpy(), rng_cpu=torch.get_rng_state().numpy(),
                 rng_device=torch.cuda.get_rng_state(self.block_model._dev).numpy())
        d.update(extra)
        tmp = path + ".tmp"
        with open(tmp, "wb") as f:
            np.savez(f, **d)
        os.replace(tmp, path)

    def load_checkpoint(self, path):
        """Restore what save_checkpoint wrote; returns all stored arrays."""
        with np.load(path, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        self.block_model.load_state_dict({k[len("block_"):]: v for k, v in d.items() if k.startswith("block_")})
        self._hp_log = torch.from_numpy(d["hp_log"].copy())
        self._hp = torch.from_numpy(d["hp"].copy())
        self._hidden_state = torch.from_numpy(d["hidden_state"].copy()).to(self.block_model._dev)
        torch.set_rng_state(torch.from_numpy(d["rng_cpu"].copy()))
        torch.cuda.set_rng_state(torch.from_numpy(d["rng_device"].copy()), self.block_model._dev)
        return d

    def da_rest_run(self, bold_real, write_path, observation_times=None, steps=800, bold_sigma=1e-8,
                    solo_rate=0.8, hp_sigma=0.25, save=True, checkpoint=None):
        """data_assimilation/DataAssimilation.py:402-453: one analysis per window.

        `bold_real` is [T, voxels per member] or a callable(t) -> [voxels per member]
        (an observation produced while the windows run).  Returns the per-window
        hidden states before and after the analysis and the window times.

        With `checkpoint` (a file path) the run is restartable: the state is
        saved there after every window, and a run that finds the file resumes
        after its last window with the same results as an uninterrupted run
        (bit-for-bit when delivery sums are order-independent, i.e. bucketed
        delivery with fixed-point accumulation)."""
        T = (bold_real.shape[0] if observation_times is None else observation_times) if not callable(bold_real) \
            else observation_times
        w_save, w_fix = [], []
        start = 0
        if checkpoint is not None and os.path.exists(checkpoint):
            d = self.load_checkpoint(checkpoint)
            start = int(d["window"])
            w_save, w_fix = list(d["w_save"]), list(d["w_fix"])
        for t in range(start, T):
            t0 = time.perf_counter()
            self.da_evolve(steps, hp_sigma=hp_sigma)
            w_save.append(self._hidden_state.cpu().numpy())
            y = bold_real(t) if callable(bold_real) else bold_real[t]
            self.da_filter(torch.as_tensor(y).reshape(1, -1), bold_sigma=bold_sigma, solo_rate=solo_rate)
            w_fix.append(self._hidden_state.cpu().numpy())
            torch.cuda.synchronize()
            self.timings.append(time.perf_counter() - t0)
            if checkpoint is not None:
                self.save_checkpoint(checkpoint, window=np.int64(t + 1), w_save=np.stack(w_save),
                                     w_fix=np.stack(w_fix))
        w_save = np.stack(w_save)
        if save:
            os.makedirs(write_path, exist_ok=True)
            np.save(os.path.join(write_path, "w.npy"), w_save)
            np.save(os.path.join(write_path, "w_fix.npy"), np.stack(w_fix))
            np.save(os.path.join(write_path, "bold_assimilation.npy"), w_save[:, :, :, -1])
        return w_save, np.stack(w_fix), np.array(self.timings)
