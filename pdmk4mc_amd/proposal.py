"""Training and judging of the learned proposal (no reference counterpart).

The proposal is the MLP of hpdmk_mc_set_proposal: delta = mu(phi(o)) +
sigma(phi(o)) z.  Its quality is the expected squared jump distance of the
Metropolis-Hastings chain it drives,

    J = E_{o ~ pi, delta ~ q(.|o)} [ alpha(o, delta) |delta|^2 ],
    alpha = min(1, exp(-beta dE + H)),  H = log q(-delta | phi(t)) - log q(delta | phi(o)),

and the uniform proposal with its step size tuned on the same system is the
judge.  Training runs on the GPU (hpdmk_mlp_train) in two stages:

* warm start: maximum likelihood of the accepted displacements of a uniform
  chain (the chain's own configurations and origin features);
* policy gradient on J from the learned chain's own moves.  Differentiating
  J through both the sampling density and alpha (alpha depends on the
  weights through H) gives, per move,

      grad J = 1{alpha = 1} |delta|^2 grad log q(delta | phi(o))
             + 1{alpha < 1} alpha |delta|^2 grad log q(-delta | phi(t)),

  so a round trains on two sample sets with signed weights: the origin
  features with the displacement and weight 1{alpha = 1} |delta|^2 - b, and
  the target features with the reversed displacement and weight
  1{alpha < 1} alpha |delta|^2 (b = the mean reward, a baseline: the
  expectation of grad log q(delta | phi(o)) is zero).  A hard-sphere overlap
  has alpha = 0 and contributes only the baseline term.
"""
import time

import numpy as np

from .hpdmk import mlp_init, mlp_train


def acceptance_probability(dE, H, beta):
    """alpha = min(1, exp(-beta dE + H)) of each move; 0 where dE is infinite."""
    x = np.where(np.isfinite(dE), -beta * np.where(np.isfinite(dE), dE, 0.0) + H, -np.inf)
    return np.exp(np.minimum(x, 0.0))


def jump_size(delta, dE, beta, objective):
    """Squared jump of each move: |delta|^2 ("distance") or (beta dE)^2
    ("energy", 0 where dE is infinite: such a move is never accepted)."""
    if objective == "distance":
        return (delta * delta).sum(1)
    if objective == "energy":
        return np.where(np.isfinite(dE), beta * np.where(np.isfinite(dE), dE, 0.0), 0.0) ** 2
    raise ValueError("objective must be 'distance' or 'energy'")


def policy_gradient_samples(feat_o, feat_t, delta, dE, H, beta, objective="distance"):
    """Signed-weight samples (features, displacements, weights) whose weighted
    log-likelihood has the policy gradient of the mean accepted squared jump
    (jump_size) as its gradient; weights are scaled to unit standard
    deviation of the reward.  Returns (feat, delta, weight, mean reward)."""
    alpha = acceptance_probability(dE, H, beta)
    d2 = jump_size(delta, dE, beta, objective)
    reward = alpha * d2
    b = reward.mean()
    full = alpha >= 1.0
    w_o = np.where(full, d2, 0.0) - b
    part = (~full) & (alpha > 0.0)
    w_t = reward[part]
    scale = 1.0 / max(reward.std(), 1e-12)
    feat = np.concatenate([feat_o, feat_t[part]])
    dl = np.concatenate([delta, -delta[part]])
    w = np.concatenate([w_o, w_t]) * scale
    return feat, dl, w, b


def run_uniform(tree, n_particles, step, moves, beta, batch, rng, rf=0.0):
    """A batched chain of `moves` uniform moves in [-step, step]^3 on tree's
    current state; returns (dE, accepted, displacements, features, seconds)."""
    tree.mc_set_proposal(False, rf=rf)
    idx = rng.integers(0, n_particles, moves)
    d = rng.uniform(-step, step, (moves, 3))
    u = rng.uniform(0.0, 1.0, moves)
    tree.mc_load(idx, d, u)
    t0 = time.perf_counter()
    tree.mc_run(beta, 1, batch)
    sec = time.perf_counter() - t0
    dE, acc = tree.mc_results()
    feat = tree.mc_proposal_results()[1] if rf > 0 else None
    return dE, acc, d, feat, sec


def run_learned(tree, n_particles, weights, cfg, moves, beta, batch, rng):
    """A batched chain of `moves` learned-proposal moves on tree's current
    state; returns (dE, accepted, displacements, H, origin and target
    features, seconds)."""
    tree.mc_set_proposal(True, weights=weights, **cfg)
    idx = rng.integers(0, n_particles, moves)
    z = rng.normal(size=(moves, 3))
    u = rng.uniform(0.0, 1.0, moves)
    tree.mc_load(idx, z, u)
    t0 = time.perf_counter()
    tree.mc_run(beta, 1, batch)
    sec = time.perf_counter() - t0
    dE, acc = tree.mc_results()
    dl, ft, H = tree.mc_proposal_results()
    ftt = tree.mc_target_features()
    return dE, acc, dl, H, ft, ftt, sec


def esjd(dE, H, delta, beta, objective="distance"):
    """Expected squared jump per move: mean of alpha jump_size (the
    acceptance probability in place of the realised decision)."""
    return float((acceptance_probability(dE, H, beta) * jump_size(delta, dE, beta, objective)).mean())


def tune_uniform(tree, n_particles, steps, moves, beta=1.0, batch=64, seed=0):
    """The judge: the uniform step size with the highest expected squared
    jump per second on tree's system (the chain continues on tree).  Returns
    (best step, {step: dict of acceptance, esjd_per_move, moves_per_s,
    esjd_per_s})."""
    rng = np.random.default_rng(seed)
    out = {}
    for s in steps:
        dE, acc, d, _, sec = run_uniform(tree, n_particles, s, moves, beta, batch, rng)
        e = esjd(dE, np.zeros(len(dE)), d, beta)
        out[float(s)] = {"acceptance": float(acc.mean()), "esjd_per_move": e, "moves_per_s": moves / sec,
                         "esjd_per_s": e * moves / sec}
    best = max(out, key=lambda s: out[s]["esjd_per_s"])
    return best, out


def train_proposal(tree, n_particles, cfg, beta=1.0, batch=64, seed=0, warm_step=1.0, warm_moves=100000,
                   rounds=10, round_moves=100000, lr=3e-4, epochs=1, pg_batch=16384, warm_epochs=20, weights0=None,
                   objective="distance"):
    """Trains the proposal MLP on tree's own chain (which continues on tree):
    maximum likelihood of the accepted moves of `warm_moves` uniform moves of
    step warm_step (skipped when weights0 is given), then `rounds` policy
    gradient rounds on J, each on `round_moves` moves of the current learned
    chain and a few large-minibatch Adam steps (`epochs` passes of minibatches
    of `pg_batch` samples: the estimate is on-policy, so each round moves the
    weights only a little).  cfg: rf, max_mean, sigma_min, sigma_max.  Returns (weights with
    the highest measured J, history of per-round dicts)."""
    rng = np.random.default_rng(seed)
    scfg = {k: cfg[k] for k in ("max_mean", "sigma_min", "sigma_max")}
    hist = []
    if weights0 is None:
        dE, acc, d, ft, _ = run_uniform(tree, n_particles, warm_step, warm_moves, beta, batch, rng, rf=cfg["rf"])
        w, losses = mlp_train(ft, d, acc.astype(np.float64), mlp_init(seed), lr=3e-3, epochs=warm_epochs,
                              batch=1024, seed=seed, **scfg)
        hist.append({"stage": "warm", "moves": warm_moves, "acceptance": float(acc.mean()),
                     "nll_first": float(losses[0]), "nll_last": float(losses[-1])})
    else:
        w = np.array(weights0, dtype=np.float32, copy=True)
    best_w, best_j = w, -1.0
    for k in range(rounds + 1):
        dE, acc, dl, H, ft, ftt, sec = run_learned(tree, n_particles, w, cfg, round_moves, beta, batch, rng)
        j = esjd(dE, H, dl, beta, objective)
        hist.append({"stage": "round", "round": k, "objective_per_move": j, "esjd_per_move": esjd(dE, H, dl, beta),
                     "acceptance": float(acc.mean()),
                     "moves_per_s": round_moves / sec, "mean_sigma_proxy": float(np.sqrt((dl * dl).mean()))})
        if j > best_j:
            best_w, best_j = w, j
        if k == rounds:
            break
        f, d, wt, _ = policy_gradient_samples(ft, ftt, dl, dE, H, beta, objective)
        w, _ = mlp_train(f, d, wt, w, lr=lr, epochs=epochs, batch=pg_batch, seed=seed + k + 1, signed=True, **scfg)
    return best_w, hist
