"""Frontend dispatch by class swap, as kymatio/frontend/entry.py:6-47 does.

Without a frontend= argument the NumPy frontend is chosen, as the
reference's entry classes do (kymatio/frontend/entry.py:20-21).  The 'numpy'
(arrays staged through the device), 'torch' and 'sklearn' (transformers over
the numpy frontend) frontends exist; other frontend names raise the
reference's RuntimeError.
"""
import importlib
import logging

_SUFFIX = {'torch': 'Torch', 'numpy': 'NumPy', 'sklearn': 'Transformer'}
_KNOWN = ['torch', 'numpy', 'tensorflow', 'jax', 'keras', 'sklearn']


class ScatteringEntry:
    def __init__(self, *args, **kwargs):
        self.name = kwargs.pop('name')
        self.class_name = kwargs.pop('class_name')
        frontend = kwargs.pop('frontend', 'numpy').lower()
        if frontend not in _KNOWN:
            raise RuntimeError('The frontend \'%s\" is not valid. Must be one of \'%s\', or \'%s\'.'
                               % (frontend, '\', \''.join(_KNOWN[:-1]), _KNOWN[-1]))
        if frontend not in _SUFFIX:
            raise RuntimeError('The frontend \'%s\' is not provided by kymatio_amd (torch, numpy, sklearn).' % frontend)
        module = importlib.import_module('kymatio_amd.' + self.class_name + '.frontend.' + frontend + '_frontend')
        cls = getattr(module, self.__class__.__name__.replace('Entry', _SUFFIX[frontend]))
        self.__class__ = cls
        self.__init__(*args, **kwargs)
        logging.info('The ' + self.name + ' frontend ' + frontend + ' was imported.')
