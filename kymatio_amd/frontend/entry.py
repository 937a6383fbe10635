"""Frontend dispatch by class swap, as kymatio/frontend/entry.py:6-47 does.

Only the 'torch' frontend exists in this package (the MI355X path); it is
also the default.  Other frontend names raise the reference's RuntimeError.
"""
import importlib
import logging

_SUFFIX = {'torch': 'Torch'}
_KNOWN = ['torch', 'numpy', 'tensorflow', 'jax', 'keras', 'sklearn']


class ScatteringEntry:
    def __init__(self, *args, **kwargs):
        self.name = kwargs.pop('name')
        self.class_name = kwargs.pop('class_name')
        frontend = kwargs.pop('frontend', 'torch').lower()
        if frontend not in _KNOWN:
            raise RuntimeError('The frontend \'%s\" is not valid. Must be one of \'%s\', or \'%s\'.'
                               % (frontend, '\', \''.join(_KNOWN[:-1]), _KNOWN[-1]))
        if frontend not in _SUFFIX:
            raise RuntimeError('The frontend \'%s\' is not provided by kymatio_amd (torch only).' % frontend)
        module = importlib.import_module('kymatio_amd.' + self.class_name + '.frontend.' + frontend + '_frontend')
        cls = getattr(module, self.__class__.__name__.replace('Entry', _SUFFIX[frontend]))
        self.__class__ = cls
        self.__init__(*args, **kwargs)
        logging.info('The ' + self.name + ' frontend ' + frontend + ' was imported.')
