from ..frontend.entry import ScatteringEntry


class ScatteringEntry2D(ScatteringEntry):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, name='Scattering2D', class_name='scattering2d', **kwargs)


__all__ = ['ScatteringEntry2D']
