"""Python algorithm plugins (python/astra/plugin.py, python/astra/plugin_c.pyx;
registry semantics of python/astra/src/PythonPluginAlgorithmFactory.cpp:108-230,
create / run of python/astra/src/PythonPluginAlgorithm.cpp:98-125).

A plugin is a class with an ``astra_name``, an ``initialize(self, cfg,
<required options>, <options with defaults>)`` and a ``run(self,
iterations)``.  Once registered, its name is an algorithm type:
``algorithm.create`` instantiates the class and calls ``astra_init(cfg)``
(which, for subclasses of ``base``, checks the configuration's options
against initialize's signature and calls it), ``algorithm.run`` calls
``run(iterations)`` and ``algorithm.get_plugin_object`` returns the
instance.  Plugins run in Python; the projections they take (``OpTomo``,
``create_sino``, the ``experimental`` calls) run on the GPU.
"""
import importlib
import inspect

from . import data2d, data3d, log, projector
from .log import AstraError

__all__ = ["base", "ReconstructionAlgorithm2D", "ReconstructionAlgorithm3D", "register",
           "get_registered", "get_help"]

# name -> plugin class, or "module.Class" (imported when an algorithm is made)
_plugins = {}


class base:
    """Base class of plugins: maps the configuration's options onto the
    arguments of ``initialize``."""

    def astra_init(self, cfg):
        spec = inspect.getfullargspec(self.initialize)
        names = spec.args[2:]  # after self and cfg
        nopt = len(spec.defaults or ())
        req, opt = names[:len(names) - nopt], names[len(names) - nopt:]
        given = dict(cfg.get("options") or cfg.get("option") or {})
        missing = [k for k in req if k not in given]
        for k in missing:
            log.error("Required option '%s' for plugin '%s' not specified"
                      % (k, type(self).__name__))
        if missing:
            raise AstraError("Missing required options")
        unused = sorted(set(given) - set(req) - set(opt))
        if unused:
            log.warn("%s: unused configuration option: %s" % (type(self).__name__, unused))
        self.initialize(cfg, *[given[k] for k in req], **{k: given[k] for k in opt if k in given})


class ReconstructionAlgorithm2D(base):
    """2D plugin base: sets ``pid``, ``s`` / ``v`` (numpy views of the
    projection and reconstruction data) and ``vg`` / ``pg``."""

    def astra_init(self, cfg):
        from ._registry import projectors2d
        self.pid = cfg["ProjectorId"]
        self.s = data2d.get_shared(cfg["ProjectionDataId"])
        self.v = data2d.get_shared(cfg["ReconstructionDataId"])
        self.vg = projector.volume_geometry(self.pid)
        self.pg = projector.projection_geometry(self.pid)
        p = projectors2d.get(self.pid)
        if self.s.shape != tuple(p.proj_shape):
            raise AstraError("Projection data and projector not compatible")
        if self.v.shape != tuple(p.vol_shape):
            raise AstraError("Reconstruction data and projector not compatible")
        super().astra_init(cfg)


class ReconstructionAlgorithm3D(base):
    """3D plugin base: as the 2D one, geometries from the data objects."""

    def astra_init(self, cfg):
        self.pid = cfg["ProjectorId"]
        self.s = data3d.get_shared(cfg["ProjectionDataId"])
        self.v = data3d.get_shared(cfg["ReconstructionDataId"])
        self.vg = data3d.get_geometry(cfg["ReconstructionDataId"])
        self.pg = data3d.get_geometry(cfg["ProjectionDataId"])
        super().astra_init(cfg)


def _class_from_string(path):
    mod, _, name = path.rpartition(".")
    if not mod:
        raise AstraError("plugin class '%s' must be given as module.Class" % path)
    try:
        return getattr(importlib.import_module(mod), name)
    except (ImportError, AttributeError) as e:
        raise AstraError("cannot load plugin class '%s': %s" % (path, e)) from None


def register(className, name=None):
    """Register a plugin class (a class object or "module.Class") under
    ``name``, by default the class's ``astra_name``.  A string with a name
    is imported only when an algorithm of that type is created."""
    if isinstance(className, str) and name is not None:
        _plugins[str(name)] = className
        return
    cls = _class_from_string(className) if isinstance(className, str) else className
    if name is None:
        name = getattr(cls, "astra_name", None)
        if name is None:
            log.error("plugin class %r has no astra_name" % (cls,))
            return
    _plugins[str(name)] = cls


def get_registered():
    """Registered plugins: name -> class (or "module.Class")."""
    return dict(_plugins)


def get_help(name):
    """The plugin class's docstring ("" and an error message if ``name``
    is not registered)."""
    c = _plugins.get(name)
    if c is None:
        log.error("Plugin %s not found!" % name)
        return ""
    if isinstance(c, str):
        c = _class_from_string(c)
    return inspect.getdoc(c) or ""


def lookup(name):
    """The class registered under ``name`` (imported if given as a
    string), or None."""
    c = _plugins.get(name)
    if isinstance(c, str):
        c = _class_from_string(c)
    return c
