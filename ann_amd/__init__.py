"""ann_amd: MI355X-native Behler-Parrinello energy / force / weight-gradient path
behind the AtomicNN (markdellostritto/ann) NNP interface."""
from .model import Model, Network, CUT, RAD, ANG, NEURON, pair_index, init_network, precondition  # noqa: F401
from . import configs  # noqa: F401
