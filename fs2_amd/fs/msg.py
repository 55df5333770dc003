"""Log levels, py/fs/msg.py."""
from . import _fs as c

dict_loglevel = {'debug': 0, 'verbose': 1, 'info': 2, 'warn': 3, 'error': 4,
                 'fatal': 5}


def set_loglevel(loglevel):
    """Set the amount of standard output messages (int or name)."""
    if isinstance(loglevel, str):
        loglevel = dict_loglevel[loglevel]
    c.set_loglevel(loglevel)
