from ..frontend.entry import ScatteringEntry


class ScatteringEntry1D(ScatteringEntry):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, name='Scattering1D', class_name='scattering1d', **kwargs)


__all__ = ['ScatteringEntry1D']
