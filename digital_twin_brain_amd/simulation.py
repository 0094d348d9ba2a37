"""In-process `simulation` (simulation/simulation.py:22-428) on a local block.

Same constructor and methods as the reference class; the block is a
`LocalBlockWrapper` instead of a gRPC client, and the per-step voxel merge
and BOLD update of `evolve` (simulation/simulation.py:269-279) run on the GPU
inside the step batches (dtb_block_attach_bold) instead of in Python each
step.  `self.bold` exposes the BOLD state (s, q, v, f_in) and
`state_update` like models/bold_model_pytorch.py.
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch

from .wrapper import LocalBlockWrapper

# default_params.py:11-12
bold_params = dict(epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02)
v_th = -50


class AttachedBOLD:
    """BOLD state living next to the block on the GPU (FP64), with the
    attribute surface of models/bold_model_pytorch.BOLD."""

    def __init__(self, nb, device):
        self.nb = nb
        self.device = device

    def _get(self, which):
        return torch.from_numpy(self.nb.bold_state(which)).to(self.device)

    @property
    def s(self):
        return self._get("s")

    @property
    def q(self):
        return self._get("q")

    @property
    def v(self):
        return self._get("v")

    @property
    def f_in(self):
        return self._get("f_in")

    @property
    def bold(self):
        return self._get("bold")

    def state_update(self, w):
        """models/bold_model_pytorch.py:34-41: w[..., 0:4] = (s, q, v, f_in); q, v, f_in floored at 1e-5."""
        if isinstance(w, np.ndarray):
            w = torch.from_numpy(w.astype(np.float64))
        w = w.detach().to(torch.float64).cpu()
        self.nb.set_bold_state("s", w[:, :, 0].reshape(-1).numpy())
        self.nb.set_bold_state("q", torch.clamp_min(w[:, :, 1], 1e-5).reshape(-1).numpy())
        self.nb.set_bold_state("v", torch.clamp_min(w[:, :, 2], 1e-5).reshape(-1).numpy())
        self.nb.set_bold_state("f_in", torch.clamp_min(w[:, :, 3], 1e-5).reshape(-1).numpy())


class simulation:
    """simulation/simulation.py:22 with an in-process GPU block."""

    @staticmethod
    def merge_data(a, weights=None, bins=10, range=None):
        return np.histogram(a, weights=weights, bins=bins, range=range)[0]

    def __init__(self, ip=None, block_path=None, dt=1.0, route_path=None, column=True, network=None, device=0,
                 bold_steps=1, **kwargs):
        ppv = 10 if column else 2
        self.block_model = LocalBlockWrapper(ip, block_path, dt, route_path=route_path, overlap=ppv, network=network,
                                             device=device)
        self.populations_per_voxel = ppv
        self.population_id = self.block_model.subblk_id
        self.population_id_cpu = self.population_id.cpu().numpy()
        pmax = int(self.population_id_cpu.max())
        self.num_voxels = int((pmax + 9) // 10) if column else int((pmax + 1) // 2)
        self.num_populations = int(self.block_model.total_subblks)
        self.num_neurons = int(self.block_model.total_neurons)
        self.num_neurons_per_population = self.block_model.neurons_per_subblk
        self.num_neurons_per_population_cpu = self.num_neurons_per_population.cpu().numpy()
        self.num_neurons_per_voxel_cpu = self.merge_data(self.population_id_cpu, self.num_neurons_per_population_cpu,
                                                         self.num_voxels, (0, self.num_voxels * ppv))
        # voxel of each sub-population exactly as np.histogram bins it (last edge inclusive)
        nb = self.block_model.nb
        vop = np.full(nb.S, -1, np.int32)
        hi = self.num_voxels * ppv
        for p in self.population_id_cpu:
            if 0 <= p < hi:
                vop[p] = p // ppv
            elif p == hi:
                vop[p] = self.num_voxels - 1
        self._vop = vop
        # BOLD advances delta_t = 1 ms per call (models/bold_model_pytorch.py:13); the reference
        # calls it once per network step (simulation/simulation.py:279)
        self.bold_steps = int(bold_steps)
        nb.attach_bold(**bold_params, delta_t=1e-3, steps_per_sample=self.bold_steps, voxel_of_pop=vop,
                       voxel_size=self.num_neurons_per_voxel_cpu.astype(np.float64), nvox=self.num_voxels)
        self.bold = AttachedBOLD(nb, self.block_model._dev)
        self.write_path = kwargs.get("write_path", "./")
        os.makedirs(self.write_path, exist_ok=True)
        self.name = kwargs.get("name", "simulation")
        self.print_info = kwargs.get("print_info", False)
        self.vmean_option = kwargs.get("vmean_option", False)
        self.sample_option = kwargs.get("sample_option", False)
        self.imean_option = kwargs.get("imean_option", False)

    def clear_mode(self):
        self.sample_option = False
        self.vmean_option = False
        self.imean_option = False
        self.print_info = False

    def update(self, hp_index, param):
        """simulation/simulation.py:122-144: Gamma(param*1e8, 1e8) ~ a constant."""
        info = torch.stack(torch.meshgrid(self.population_id.cpu(), torch.tensor([hp_index], dtype=torch.int64),
                                          indexing="ij"), dim=-1).reshape((-1, 2))
        alpha = torch.ones(self.num_populations) * param * 1e8
        beta = torch.ones(self.num_populations) * 1e8
        self.block_model.gamma_property_by_subblk(info, alpha, beta, debug=False)

    def sample(self, sample_idx):
        sample_idx = torch.as_tensor(sample_idx, dtype=torch.int64)
        assert int(sample_idx.max()) < self.num_neurons
        self.block_model.set_samples(sample_idx)
        self.num_sample = int(sample_idx.shape[0])
        return self.num_sample

    def mul_property_by_subblk(self, index, value):
        assert isinstance(index, torch.Tensor) and isinstance(value, torch.Tensor)
        self.block_model.mul_property_by_subblk(index, value)

    def gamma_initialize(self, hp_index, population_id=None, alpha=5.0, beta=5.0):
        """simulation/simulation.py:195-223."""
        population_id = self.population_id if population_id is None else population_id
        info = torch.stack(torch.meshgrid(population_id.cpu(), torch.tensor([hp_index], dtype=torch.int64),
                                          indexing="ij"), dim=-1).reshape((-1, 2))
        a = torch.ones(len(population_id)) * alpha
        b = torch.ones(len(population_id)) * beta
        return self.block_model.gamma_property_by_subblk(info, a, b, debug=False)

    def _act(self, freqs_last):
        f = freqs_last.detach().cpu().numpy()
        act = self.merge_data(self.population_id_cpu, f, self.num_voxels,
                              (0, self.num_voxels * self.populations_per_voxel))
        return torch.from_numpy((act / self.num_neurons_per_voxel_cpu).reshape(-1)).to(self.block_model._dev)

    def evolve(self, step, vmean_option=False, sample_option=False, imean_option=False, bold_detail=False,
               window_input=False):
        """simulation/simulation.py:225-299: `step` network steps; BOLD per step on the GPU.

        window_input (with bold_detail): also keep, in `last_window_input` [voxels], the
        window's mean BOLD input (spikes of the voxel per BOLD step over its neurons)."""
        nb = self.block_model.nb
        dev = self.block_model._dev
        S = nb.S
        sel = torch.from_numpy(self.population_id_cpu).to(dev)
        if bold_detail and not (vmean_option or sample_option or imean_option) and step > self.bold_steps:
            # the hidden state needs only the last step's counts: run the window
            # without per-step outputs, then its last BOLD sample with counts
            head = step - self.bold_steps
            tot = None
            if window_input:
                # per-population spike totals over the window, batch by batch on the device
                tot = torch.zeros(S, dtype=torch.int64, device=dev)
                chunk = 1000 - 1000 % self.bold_steps
                buf = torch.empty((chunk, S), dtype=torch.int32, device=dev)
                done = 0
                while done < head:
                    k = min(chunk, head - done)
                    nb.run(k, noise_mode=1, graph=True, freqs=buf[:k])
                    tot += buf[:k].sum(dim=0, dtype=torch.int64)
                    done += k
            else:
                nb.run(head, noise_mode=1, graph=True)
            f = torch.empty((self.bold_steps, S), dtype=torch.int32, device=dev)
            bold = torch.empty((1, self.num_voxels), dtype=torch.float64, device=dev)
            nb.run(self.bold_steps, noise_mode=1, graph=True, freqs=f, bold=bold)
            if window_input:
                tot += f.sum(dim=0, dtype=torch.int64)
                vop = torch.from_numpy(self._vop.astype(np.int64)).to(dev)
                ok = vop >= 0
                cnt = torch.zeros(self.num_voxels, dtype=torch.float64, device=dev)
                cnt.index_add_(0, vop[ok], tot[ok].double())
                size = torch.from_numpy(self.num_neurons_per_voxel_cpu.astype(np.float64)).to(dev)
                self.last_window_input = cnt / size / (step // self.bold_steps)
            b = self.bold
            return self._act(f.index_select(1, sel)[-1]), b.s, b.q, b.v, b.f_in, bold[-1]
        f = torch.empty((step, S), dtype=torch.int32, device=dev)
        vm = torch.empty((step, S), dtype=torch.float32, device=dev) if vmean_option else None
        im = torch.empty((step, S), dtype=torch.float32, device=dev) if imean_option else None
        ns = self.block_model._sample_num
        ss = torch.empty((step, ns), dtype=torch.uint8, device=dev) if sample_option else None
        sv = torch.empty((step, ns), dtype=torch.float32, device=dev) if sample_option else None
        bold = torch.empty((step // self.bold_steps, self.num_voxels), dtype=torch.float64, device=dev)
        nb.run(step, noise_mode=1, graph=True, freqs=f, vmeans=vm, imeans=im, sample_spike=ss, sample_v=sv, bold=bold)
        freqs = f.index_select(1, sel)
        bold_out = bold[-1]
        if bold_detail:
            b = self.bold
            return self._act(freqs[-1]), b.s, b.q, b.v, b.f_in, bold_out
        out = (freqs,)
        if vmean_option:
            out += (vm.index_select(1, sel),)
        if sample_option:
            order = self.block_model._sample_order
            if order is not None:
                ss, sv = ss.index_select(1, order), sv.index_select(1, order)
            spike = ss.bool() & (torch.abs(sv - v_th) / 50 < 1e-5)  # simulation/simulation.py:292
            out += (spike, sv)
        if imean_option:
            out += (im.index_select(1, sel),)
        out += (bold_out,)
        return out

    def run(self, step=800, observation_time=100, hp_index=None, hp_total=None):
        """simulation/simulation.py:301-395 (npy outputs; .mat export omitted)."""
        start = time.time()
        if hp_total is not None:
            for k in hp_index:
                self.gamma_initialize(k, self.population_id)
            info = torch.stack(torch.meshgrid(self.population_id.cpu(), torch.tensor(hp_index, dtype=torch.int64),
                                              indexing="ij"), dim=-1).reshape((-1, 2))
            self.block_model.mul_property_by_subblk(info, hp_total[0].reshape(-1).float())
            hp_total = hp_total[1:]
            state = "after"
        else:
            state = "before"
        self.evolve(3200)
        bolds = np.zeros([observation_time, self.num_voxels], dtype=np.float32)
        freqs_all = []
        for i in range(observation_time):
            if hp_total is not None:
                self.block_model.mul_property_by_subblk(info, hp_total[i].reshape(-1).float())
            out = self.evolve(step, vmean_option=self.vmean_option, sample_option=self.sample_option,
                              imean_option=self.imean_option)
            freqs_all.append(out[0].cpu().numpy().astype(np.uint32))
            bolds[i] = out[-1].cpu().numpy()
        np.save(os.path.join(self.write_path, f"freqs_{state}_assim.npy"), np.stack(freqs_all))
        np.save(os.path.join(self.write_path, f"bold_{state}_assim.npy"), bolds)
        if self.print_info:
            print(f"done in {time.time() - start:.2f} s")
        return bolds

    def __call__(self, step=800, observation_time=100, hp_index=None, hp_path=None):
        hp_total = None
        if hp_path is not None:
            hp_total = torch.from_numpy(np.load(hp_path).astype(np.float32))
        out = self.run(step=step, observation_time=observation_time, hp_index=hp_index, hp_total=hp_total)
        self.block_model.shutdown()
        return out
