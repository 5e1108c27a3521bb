"""Learned proposal on a 128^3 Lebwohl-Lasher lattice near the transition.

Equilibrates a lattice with plain sweeps, trains the 22-64-64-4 proposal
network on neighbourhoods drawn from it, then runs learned sweeps (exact
Metropolis-Hastings with the network's axial von Mises-Fisher proposal, then
the reflection of each director through the network's mean direction) and
prints their rate and acceptance (of the vMF move) next to plain sweeps of the
same lattice.

    python mclics-backend_amd/examples/learned_proposal.py [--sweeps N] [--T T]
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import mclics  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--T", type=float, default=1.12)
    ap.add_argument("--sweeps", type=int, default=200, help="learned sweeps to time")
    ap.add_argument("--samples", type=int, default=262144, help="training neighbourhoods")
    args = ap.parse_args()
    L = args.L
    with mclics.Network(seed=5) as net, \
            mclics.Lattice(mclics.Config(L, L, L, T=args.T, seed=3), mclics.random_ic(L, L, L, seed=3)) as lat:
        lat.sweeps(500)
        for k in range(4):
            lat.sweeps(20)
            net.add_samples(lat, args.samples // 4, seed=k)
        losses = net.train(epochs=3, batch=2048, lr=2e-3)
        print(f"trained on {net.dataset_size()} neighbourhoods, loss per epoch {[round(float(v), 4) for v in losses]}")
        ms, _ = lat.time_sweeps(args.sweeps, graph=False)
        print(f"plain:   {args.sweeps / (ms / 1e3):9.1f} sweeps/s  acceptance {lat.acceptance()[0] / L**3:.3f}")
        lat.set_learned(net)
        lat.sweeps(5)
        ms, _ = lat.time_sweeps(args.sweeps, graph=False)
        print(f"learned: {args.sweeps / (ms / 1e3):9.1f} sweeps/s  acceptance {lat.acceptance()[0] / L**3:.3f}")


if __name__ == "__main__":
    main()
