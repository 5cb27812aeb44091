"""Keyword-only option decorators of the fractal API.

``zoom_options`` stores the call's keyword arguments (defaults applied) as
instance attributes and in ``self.zoom_kwargs``; ``calc_options`` forwards
the arguments and the returned factory dict to ``self.calc_hook``.  Same
contract as the reference decorators (src/fractalshades/utils.py:247-335,
399-406).
"""
import functools
import inspect


def _bound_kwargs(method, args, kwargs):
    if args:
        raise TypeError(
            f"{method.__name__} should only accept keyword-arguments ; "
            f"given positionnal: {args}"
        )
    ba = inspect.signature(method).bind_partial(**kwargs)
    ba.apply_defaults()
    return dict(ba.arguments)


def zoom_options(method):
    @functools.wraps(method)
    def wrapper(self, *args, **kwargs):
        kw = _bound_kwargs(method, args, kwargs)
        self.zoom_kwargs = kw
        for k, v in kw.items():
            setattr(self, k, v)
        return method(self, **kwargs)
    wrapper._zoom_options = True
    return wrapper


def calc_options(method):
    @functools.wraps(method)
    def wrapper(self, *args, **kwargs):
        kw = _bound_kwargs(method, args, kwargs)
        return_dic = method(self, **kwargs)
        self.calc_hook(method.__name__, kw, return_dic)
        return return_dic
    wrapper._calc_options = True
    return wrapper
