"""Parameter attribute bags (reference: helpers/parameters.py:86-112).

A ``Parameters`` subclass sets its parameters as instance attributes in
``__init__``; ``check()`` reports those still None.
"""


class TreePrintable:
    """Readable repr: class name followed by the public attributes, one per line."""

    def __repr__(self):
        attrs = sorted((k, v) for k, v in vars(self).items() if not k.startswith("_"))
        body = "\n".join(f"    {k}: {v!r}" for k, v in attrs) or "    [no attributes]"
        return f"{type(self).__name__}\n{body}"


class Parameters(TreePrintable):

    def check(self, raise_e=False):
        """Return the names of unset (None) parameters; raise RuntimeError
        instead if raise_e is True and any is unset."""
        unset = sorted(k for k, v in vars(self).items() if v is None)
        if unset:
            msg = "The following parameters have not been set:\n" + "\n".join(unset)
            if raise_e:
                raise RuntimeError(msg)
            print(msg)
            return unset
        return []
