"""Empty, as py/fs/functions.py is; present so that `fs.functions` imports."""
