"""Learned control variate for orientation averages.

An orientation average estimates E_o[M(o)] by sum_o w_o M(o) over N sampled
orientations.  With a cheap predictor g(o) of the (azimuth-averaged) Mueller
field,
    M_cv = sum_o w_o (M(o) - c g(o)) + c E[g]        (sum_o w_o = 1)
is unbiased for every fixed g and c, and has lower variance where g tracks
M.  Here g is a multilayer perceptron (mbs_mlp_*, trained on the device by
the library's own BF16-MFMA kernels) that maps the particle orientation
(unit quaternion, qw >= 0) and log size to the Mueller field on the grid's
theta rows, learned from Mueller matrices the engine itself computed
(mbs_diffract with MBS_PER_ORIENT, azimuth-averaged on the device).  E[g] is
taken over a dense device-generated Sobol orientation set of the same
symmetry domain; c is fitted per (theta, element) on the training data
(Cov(M, g) / Var(g), clipped to [0, 1.5]) and then fixed, so the estimate
stays unbiased.  The plain sum remains available (c = 0) and is the judge of
the control variate's error.

Targets are log P11 and P_ij / P11 per theta row, standardised per output;
predictions are mapped back to linear Mueller elements before use.
"""
import numpy as np

from . import abi, engine


def features(orient_rows, log_size):
    """[n, 5] float32 inputs from orientation rows (mbs_orientation_set
    layout: quaternion in columns 3..6) and the log particle size."""
    q = np.asarray(orient_rows, dtype=np.float64)[:, abi.ORIENT_Q:abi.ORIENT_Q + 4]
    q = q * np.where(q[:, 3:4] < 0, -1.0, 1.0)
    ls = np.broadcast_to(np.asarray(log_size, dtype=np.float64), (len(q),))
    return np.concatenate([q, ls[:, None]], 1).astype(np.float32)


def unweighted_rows(rows, weights):
    """Per-orientation azimuth-averaged Mueller rows [n, n_theta, 16] with
    each orientation's weight divided out."""
    rows = np.asarray(rows, dtype=np.float64)
    return rows[:, :, 2:] / np.asarray(weights, dtype=np.float64)[:, None, None]


class ControlVariate:
    def __init__(self, ctx, n_theta, width=256, depth=4, seed=0):
        self.ctx = ctx
        self.n_theta = n_theta
        self.net = engine.Mlp(ctx, 5, width, depth, 16 * n_theta, seed=seed)
        self.mu = self.sd = None
        self.c = None

    def close(self):
        self.net.close()

    def _encode(self, M):
        p11 = np.maximum(M[..., 0], 1e-300)
        t = np.concatenate([np.log(p11)[..., None], M[..., 1:] / p11[..., None]], -1)
        return t.reshape(len(M), -1)

    def predict(self, x, chunk=8192):
        """Linear Mueller prediction [n, n_theta, 16]."""
        out = []
        for i in range(0, len(x), chunk):
            t = self.net.forward(x[i:i + chunk]).astype(np.float64) * self.sd + self.mu
            t = t.reshape(-1, self.n_theta, 16)
            p11 = np.exp(t[..., 0])
            out.append(np.concatenate([p11[..., None], t[..., 1:] * p11[..., None]], -1))
        return np.concatenate(out) if out else np.zeros((0, self.n_theta, 16))

    def fit(self, x, M, epochs=200, batch=128, lr=1e-3, seed=0):
        """Trains g on inputs x [n, 5] and unweighted Mueller rows M [n,
        n_theta, 16]; fits the coefficient c.  Returns the final loss."""
        t = self._encode(M)
        self.mu, self.sd = t.mean(0), t.std(0) + 1e-6
        loss = self.net.train(x, ((t - self.mu) / self.sd).astype(np.float32), batch=batch,
                              epochs=epochs, lr=lr, seed=seed)
        g = self.predict(x)
        cov = np.mean((M - M.mean(0)) * (g - g.mean(0)), 0)
        var = np.var(g, 0)
        self.c = np.clip(np.where(var > 0, cov / np.where(var > 0, var, 1.0), 0.0), 0.0, 1.5)
        return loss

    def expectation(self, kind, n, seed, beta_sym, gamma_sym, log_size):
        """E[g] over n device-generated orientations of the domain."""
        rows = self.ctx.orientation_set(kind, n, seed=seed, beta_sym=beta_sym,
                                        gamma_sym=gamma_sym)
        return self.predict(features(rows, log_size)).mean(0)

    def estimate(self, M, x, weights, Eg, c=None):
        """sum_o w_o (M_o - c g_o) + c E[g] (weights normalised to sum 1)."""
        w = np.asarray(weights, dtype=np.float64)
        w = w / w.sum()
        c = self.c if c is None else c
        g = self.predict(x)
        return np.tensordot(w, M - c * g, axes=1) + c * Eg
