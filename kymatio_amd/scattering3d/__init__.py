from ..frontend.entry import ScatteringEntry


class HarmonicScatteringEntry3D(ScatteringEntry):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, name='HarmonicScattering3D', class_name='scattering3d', **kwargs)


__all__ = ['HarmonicScatteringEntry3D']
