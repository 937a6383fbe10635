"""torch.nn.Module base of the scattering frontends
(kymatio/frontend/torch_frontend.py:4-18): forward = input checks + scattering."""
import torch.nn as nn


class ScatteringTorch(nn.Module):
    def __init__(self):
        super().__init__()
        self.frontend_name = 'torch'

    def register_filters(self):
        raise NotImplementedError

    def forward(self, x):
        self.backend.input_checks(x)
        return self.scattering(x)
