"""Parameter tree with the behaviour of fluiddyn's ParamContainer as fluidsim
uses it (lib/fluidsim_core/params.py:22-78): children are created with
``_set_child``, attributes with ``_set_attribs``/``_set_attrib``; setting an
attribute that was not declared raises AttributeError (typo guard).
"""

import copy
import json


class Parameters:
    def __init__(self, tag="params", attribs=None, doc=""):
        object.__setattr__(self, "_tag", tag)
        object.__setattr__(self, "_key_attribs", [])
        object.__setattr__(self, "_tag_children", [])
        object.__setattr__(self, "_doc", doc)
        if attribs:
            self._set_attribs(attribs)

    # --- construction ---------------------------------------------------
    def _set_attrib(self, key, value):
        if key not in self._key_attribs:
            self._key_attribs.append(key)
        object.__setattr__(self, key, value)

    def _set_attribs(self, attribs):
        for k, v in dict(attribs).items():
            self._set_attrib(k, v)

    def _set_child(self, tag, attribs=None, doc=""):
        if tag in self._tag_children:
            child = getattr(self, tag)
            if attribs:
                child._set_attribs(attribs)
            return child
        child = Parameters(tag=tag, attribs=attribs, doc=doc)
        self._tag_children.append(tag)
        object.__setattr__(self, tag, child)
        return child

    def _set_doc(self, doc):
        object.__setattr__(self, "_doc", doc)

    def _get_formatted_docs(self):
        return self._doc

    # --- access -----------------------------------------------------------
    def __setattr__(self, key, value):
        if key in self._key_attribs:
            object.__setattr__(self, key, value)
        elif key in self._tag_children:
            raise AttributeError(f"{key} is a child of {self._tag}, not an attribute")
        else:
            raise AttributeError(
                f'"{key}" is not a declared attribute of "{self._tag}" '
                f"(declared: {self._key_attribs})"
            )

    def __getitem__(self, key):
        return getattr(self, key)

    def __setitem__(self, key, value):
        setattr(self, key, value)

    def __contains__(self, key):
        return key in self._key_attribs or key in self._tag_children

    def _make_dict_tree(self):
        d = {k: getattr(self, k) for k in self._key_attribs}
        for t in self._tag_children:
            d[t] = getattr(self, t)._make_dict_tree()
        return d

    def _update_from_dict_tree(self, tree):
        """Set attributes and children from a tree of _make_dict_tree (dict
        values are children; attributes and children missing here are
        created)."""
        for k, v in tree.items():
            if isinstance(v, dict):
                self._set_child(k)._update_from_dict_tree(v)
            elif k in self._tag_children:
                raise ValueError(f"{k} is a child of {self._tag} in this tree")
            else:
                self._set_attrib(k, v)
        return self

    def __deepcopy__(self, memo):
        new = Parameters(tag=self._tag, doc=self._doc)
        for k in self._key_attribs:
            new._set_attrib(k, copy.deepcopy(getattr(self, k), memo))
        for t in self._tag_children:
            new._tag_children.append(t)
            object.__setattr__(new, t, copy.deepcopy(getattr(self, t), memo))
        return new

    def __repr__(self):
        return f"<Parameters {self._tag}>\n" + json.dumps(
            self._make_dict_tree(), indent=1, default=str
        )

    __str__ = __repr__
