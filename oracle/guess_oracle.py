"""FP64 numpy restatement of the learned SCC start-charge model (test infrastructure).

Only tests and the bench's checks use this module; the product runs the HIP
kernels in gfn1-rs_amd/csrc/guess.hip.  The reference has no learned guess: its
start vector is zero or caller-supplied (src/electronic.rs:386-392, 1231-1239),
and the model only chooses that vector.

Descriptor per atom (68 values): one-hot of the element over (H, C, N, O), then
for each neighbour element e and k = 0..15
    G[e, k] = sum_j exp(-eta (r_ij - mu_k)^2) * fc(r_ij),   r_ij < 5 A,
with mu_k = 0.5 + 0.3 k A, eta = 1 / (2 * 0.3^2) A^-2 and
fc(r) = (cos(pi r / 5) + 1) / 2.  Network 68 -> 64 SiLU -> 64 SiLU -> 2 on
standardised inputs (x - mean) * inv_spread.  Loss: mean squared error over
the two outputs of every atom, L = 1 / (2 N) sum_a sum_o (y_ao - t_ao)^2.

Pinned in tests/test_guess.py by finite differences of the loss (backward
pass), by translation/rotation/permutation invariance and the cutoff of the
descriptor, and by the closed form of an isolated pair.
"""
import numpy as np

BOHR_TO_ANGSTROM = 1.0 / 1.8897261246257702
SLOTS = {1: 0, 6: 1, 7: 2, 8: 3}
NFEAT, NHID, NOUT = 68, 64, 2
CUTOFF, MU0, MU_STEP = 5.0, 0.5, 0.3
ETA = 1.0 / (2.0 * MU_STEP * MU_STEP)
# parameter layout (matches the C ABI's flat vector)
OFF_W1 = 0
OFF_B1 = OFF_W1 + NHID * NFEAT
OFF_W2 = OFF_B1 + NHID
OFF_B2 = OFF_W2 + NHID * NHID
OFF_W3 = OFF_B2 + NHID
OFF_B3 = OFF_W3 + NOUT * NHID
NPARAM = OFF_B3 + NOUT


def features(numbers, pos_bohr):
    """Descriptor rows (nat x 68) of one system."""
    z = np.asarray(numbers)
    x = np.asarray(pos_bohr, dtype=np.float64).reshape(-1, 3) * BOHR_TO_ANGSTROM
    nat = len(z)
    out = np.zeros((nat, NFEAT))
    mu = MU0 + MU_STEP * np.arange(16)
    for a in range(nat):
        sa = SLOTS.get(int(z[a]), -1)
        if sa >= 0:
            out[a, sa] = 1.0
        for j in range(nat):
            if j == a:
                continue
            e = SLOTS.get(int(z[j]), -1)
            if e < 0:
                continue
            r = np.linalg.norm(x[j] - x[a])
            if r >= CUTOFF:
                continue
            fc = 0.5 * (np.cos(np.pi * r / CUTOFF) + 1.0)
            out[a, 4 + 16 * e:4 + 16 * (e + 1)] += np.exp(-ETA * (r - mu) ** 2) * fc
    return out


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    w1 = t[OFF_W1:OFF_B1].reshape(NHID, NFEAT)
    b1 = t[OFF_B1:OFF_W2]
    w2 = t[OFF_W2:OFF_B2].reshape(NHID, NHID)
    b2 = t[OFF_B2:OFF_W3]
    w3 = t[OFF_W3:OFF_B3].reshape(NOUT, NHID)
    b3 = t[OFF_B3:NPARAM]
    return w1, b1, w2, b2, w3, b3


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(theta, norm, feat):
    """Network outputs (nat x 2) for standardised descriptor rows."""
    w1, b1, w2, b2, w3, b3 = unpack(theta)
    xs = (feat - norm[:NFEAT]) * norm[NFEAT:]
    z1 = xs @ w1.T + b1
    h1 = z1 * _sig(z1)
    z2 = h1 @ w2.T + b2
    h2 = z2 * _sig(z2)
    return h2 @ w3.T + b3


def loss_and_gradient(theta, norm, feat, labels):
    """L = 1/(2N) sum (y - t)^2 and dL/dtheta (flat, FP64)."""
    w1, b1, w2, b2, w3, b3 = unpack(theta)
    n = feat.shape[0]
    xs = (feat - norm[:NFEAT]) * norm[NFEAT:]
    z1 = xs @ w1.T + b1
    s1 = _sig(z1)
    h1 = z1 * s1
    z2 = h1 @ w2.T + b2
    s2 = _sig(z2)
    h2 = z2 * s2
    y = h2 @ w3.T + b3
    r = y - labels
    loss = 0.5 * np.sum(r * r) / n
    dy = r / n
    g_w3 = dy.T @ h2
    g_b3 = dy.sum(0)
    dz2 = (dy @ w3) * (s2 * (1.0 + z2 * (1.0 - s2)))
    g_w2 = dz2.T @ h1
    g_b2 = dz2.sum(0)
    dz1 = (dz2 @ w2) * (s1 * (1.0 + z1 * (1.0 - s1)))
    g_w1 = dz1.T @ xs
    g_b1 = dz1.sum(0)
    grad = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3.ravel(), g_b3])
    return loss, grad


def start_vector(y, charge):
    """Per-system start vector from per-atom outputs: shells of atom a take
    y[a, 0], y[a, 1], shifted uniformly to sum to the charge."""
    q = np.asarray(y, dtype=np.float64).reshape(-1)
    return q - (q.sum() - charge) / q.size
