"""Statistics of correlated Monte Carlo series: automated blocking and the
integrated autocorrelation time.

blocking() restates the reference's Blocker (ref include/block/blocker.cpp:15-118,
after M. Jonsson, Phys. Rev. E 98, 043304 (2018)): blocking transforms of a
2^d-long series, the M_j test statistic against 95 % chi-square quantiles, and
the variance of the mean at the first level the test accepts.
"""
import numpy as np
from scipy.stats import chi2

_Q95 = chi2.ppf(0.95, np.arange(1, 65))


def blocking(x):
    """Returns (mean, standard error, block size 2^(d-K) used)."""
    x = np.asarray(x, dtype=np.float64).copy()
    n = x.size
    if n == 0 or n & (n - 1):
        raise ValueError("blocking needs a power-of-two series length")
    d = int(np.log2(n))
    mu = x.mean()
    gam = np.zeros(d)
    s = np.zeros(d)
    for k in range(d):
        X = x - mu
        m = x.size
        gam[k] = np.dot(X[:-1], X[1:]) / m
        s[k] = np.dot(X, X) / m
        x = 0.5 * (x[0::2] + x[1::2])
    M = (gam / s) ** 2 * 2.0 ** (d - np.arange(d))
    cs = np.cumsum(M[::-1])
    k = d - 1
    while k >= 0 and not cs[k] < _Q95[k]:
        k -= 1
    K = min(d - (k + 1), d - 1)
    nK = 2 ** (d - K)
    return mu, float(np.sqrt(s[K] / nK)), nK


def iat_blocking(x):
    """tau_int from blocking: var(mean) = 2 tau sigma^2 / n."""
    x = np.asarray(x, dtype=np.float64)
    var = x.var()
    if var == 0:
        return 0.5
    _, se, _ = blocking(x)
    return max(0.5, x.size * se * se / (2.0 * var))


def iat_sokal(x, c=6.0):
    """tau_int = 1/2 + sum_t rho(t) with Sokal's self-consistent window t <= c tau."""
    x = np.asarray(x, dtype=np.float64)
    x = x - x.mean()
    n = x.size
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] == 0:
        return 0.5
    rho = acf / acf[0]
    tau = 0.5
    for t in range(1, n):
        tau += rho[t]
        if t >= c * tau:
            break
    return max(0.5, tau)


def iat_many(series, method="blocking"):
    """Mean tau_int over independent chains (rows of `series`)."""
    fn = iat_blocking if method == "blocking" else iat_sokal
    return float(np.mean([fn(s) for s in series]))


def block_means(series, block):
    """[walkers][n] -> [walkers][n // block] means of consecutive blocks."""
    x = np.asarray(series, dtype=np.float64)
    n = (x.shape[1] // block) * block
    return x[:, :n].reshape(x.shape[0], -1, block).mean(axis=2)


def iat_pooled(Y, block, var_step, c=6.0):
    """tau_int (in steps) of many independent chains from their block means.

    Y [walkers][m]: means of `block` consecutive steps of each chain;
    var_step: the variance of a single step's value.  The variance of a long
    mean is what tau measures, var(mean of L steps) = 2 tau var_step / L, and
    the block means carry it exactly: tau = block tau_Y var(Y) / var_step.

    sokal: tau_Y from the autocorrelation of Y about the mean pooled over all
      chains (no per-chain mean subtraction, whose bias 2 tau / L is the
      leading error when the chains are a few tens of tau long), averaged over
      the chains, summed with Sokal's self-consistent window t <= c tau_Y.
    blocking: tau from each chain's automated blocking of Y (its own mean),
      averaged over the chains.
    """
    Y = np.asarray(Y, dtype=np.float64)
    W, m = Y.shape
    mu = Y.mean()
    X = Y - mu
    nfft = 1 << int(np.ceil(np.log2(2 * m)))
    acf = np.zeros(m)
    for i in range(0, W, 64):
        f = np.fft.rfft(X[i:i + 64], nfft, axis=1)
        acf += np.fft.irfft(f * np.conj(f), nfft, axis=1)[:, :m].sum(axis=0)
    acf /= W * (m - np.arange(m))          # unbiased lag products
    varY = acf[0]
    if varY <= 0:
        return {"tau_int_sokal": 0.5, "tau_int_blocking": 0.5}
    rho = acf / varY
    tauY = 0.5
    for t in range(1, m):
        tauY += rho[t]
        if t >= c * tauY:
            break
    tau_s = block * max(tauY, 0.5) * varY / var_step
    taus_b = []
    for y in Y:
        if m & (m - 1) == 0 and y.var() > 0:
            _, se, _ = blocking(y)
            taus_b.append(m * block * se * se / (2.0 * var_step))
    tau_b = float(np.mean(taus_b)) if taus_b else float("nan")
    return {"tau_int_sokal": float(tau_s), "tau_int_blocking": tau_b,
            "sokal_window_steps": int(block * t)}
