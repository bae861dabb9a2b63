"""Learned T2 initial guess: training pipeline and the trained weights."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "t2_guess_h2o_ccpvdz.npz")


def load_network(path=WEIGHTS, device=0):
    """The trained network (ugx.Network) from its .npz file."""
    from .. import ugx
    if not os.path.exists(path):
        raise FileNotFoundError("no trained guess network at %s (python -m ugacc_amd.learned.train_guess)" % path)
    return ugx.Network(device).load(path)
