"""Seeded synthetic workloads of the benchmark configurations.

One definition serves the benchmark and the full-size parity tests, so both
step exactly the same networks:

- ``block``: one GPU's share of the whole-brain model, 2,000,000 neurons in
  200 voxels (lognormal sizes, 80/20 E/I each), in-degree 500 (70% within
  the voxel, 30% from a sparse lognormal voxel matrix at 5% density), OU mean
  5.0 (about 10 Hz after the initial volley), one parameter record per
  population.
- ``stress``: 500,000 neurons at in-degree 5000 (2.5e9 synapses), g_ui at
  0.02 x the 10k-neuron block's, OU mean 12, sigma 0.5: bursty synchronous
  firing near 50 Hz.
- ``ensemble``: 64 instances of 20k-200k neurons (log-uniform), in-degree
  200, per-instance g_ui uniform in 0.5-1.5 x the 10k-neuron block's, OU
  mean 3.0.
- the serving stream is ``serving.poisson_requests``.
"""
import numpy as np

from .engine import DEFAULT_G_UI, gen_uniform, gen_voxel

DT_MS = 0.1
OU_TAU = 4.0

BLOCK_NEURONS, BLOCK_VOXELS, BLOCK_DEGREE = 2_000_000, 200, 500
BLOCK_MU, BLOCK_SIGMA, BLOCK_NET_SEED, BLOCK_SIM_SEED = 5.0, 0.1, 2024, 7

STRESS_NEURONS, STRESS_DEGREE, STRESS_G_SCALE = 500_000, 5000, 0.02
STRESS_MU, STRESS_SIGMA, STRESS_NET_SEED, STRESS_SIM_SEED = 12.0, 0.5, 2024, 7

ENS_INSTANCES, ENS_MIN, ENS_MAX, ENS_DEGREE, ENS_MU, ENS_SIGMA, ENS_SEED = 64, 20_000, 200_000, 200, 3.0, 0.1, 4242
ENS_SIM_SEED = 7


def block_network(neurons=BLOCK_NEURONS, seed=BLOCK_NET_SEED):
    """The voxel block; g_AMPA differs by 1e-4 relative per population (one
    parameter record per population, as after per-population assimilation
    updates)."""
    net = gen_voxel(neurons, BLOCK_VOXELS, BLOCK_DEGREE, seed=seed)
    sub = net.props[:, 3].astype(np.int64)
    net.props[:, 10] = (net.props[:, 10] * (1.0 + 1e-4 * sub)).astype(np.float32)
    return net


def block_description(net):
    return ("one GPU's block: %d neurons, %d voxels, in-degree %d, dt %.1f ms, OU mu %.1f"
            % (net.n, BLOCK_VOXELS, BLOCK_DEGREE, DT_MS, BLOCK_MU))


def stress_network(neurons=STRESS_NEURONS, seed=STRESS_NET_SEED):
    g = (np.asarray(DEFAULT_G_UI, dtype=np.float32) * np.float32(STRESS_G_SCALE)).astype(np.float32)
    return gen_uniform(neurons, STRESS_DEGREE, seed=seed, g_ui=g)


def stress_description(net):
    return ("propagation stress: %d neurons, in-degree %d, g_ui x %.2f, dt %.1f ms, OU mu %.1f sigma %.1f"
            % (net.n, STRESS_DEGREE, STRESS_G_SCALE, DT_MS, STRESS_MU, STRESS_SIGMA))


def ensemble_specs(count=ENS_INSTANCES, seed=ENS_SEED):
    """(sizes [count], g_ui [count][4], network seeds [count])."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(ENS_MIN), np.log(ENS_MAX), count)).astype(np.int64)
    g0 = np.asarray(DEFAULT_G_UI, dtype=np.float32)
    gs = (g0[None, :] * rng.uniform(0.5, 1.5, (count, 4))).astype(np.float32)
    return sizes, gs, [seed + 1 + i for i in range(count)]


def ensemble_networks(count=ENS_INSTANCES, seed=ENS_SEED):
    sizes, gs, seeds = ensemble_specs(count, seed)
    return [gen_uniform(int(n), ENS_DEGREE, seed=s, g_ui=g) for n, g, s in zip(sizes, gs, seeds)], sizes


def ensemble_description(sizes):
    return ("%d instances, %d-%d neurons log-uniform (%d total), in-degree %d, g_ui x U(0.5,1.5), OU mu %.1f, "
            "dt %.1f ms" % (len(sizes), ENS_MIN, ENS_MAX, int(np.sum(sizes)), ENS_DEGREE, ENS_MU, DT_MS))
