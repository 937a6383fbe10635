"""Frontend base: the backend plugin loader of kymatio
(kymatio/frontend/base_frontend.py:32-50), with the same rules and messages.

A backend is a string starting with the frontend name ('torch', 'torch_hip',
...) resolved to the module ``<package>.<name>_backend``, or an object whose
``name`` starts with the frontend name.
"""
import importlib


class ScatteringBase:
    def __init__(self):
        super().__init__()

    def _instantiate_backend(self, import_string):
        if isinstance(self.backend, str):
            if not self.backend.startswith(self.frontend_name):
                raise ImportError('The backend ' + self.backend + ' can not be called from the frontend '
                                  + self.frontend_name + '.')
            try:
                self.backend = importlib.import_module(import_string + self.backend + '_backend').backend
            except ImportError:
                raise ImportError('Backend ' + self.backend + ' not found!')
        elif not getattr(self.backend, 'name', '').startswith(self.frontend_name):
            raise ImportError('The backend ' + str(getattr(self.backend, 'name', self.backend))
                              + ' is not supported by the frontend ' + self.frontend_name + '.')
