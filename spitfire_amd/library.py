"""Containers for tabulated chemistry libraries and solution trajectories (mirrors
src/spitfire/chemistry/library.py: Dimension, Library).

A Library holds named properties on the structured grid spanned by its Dimensions (meshgrid 'ij' order), with
bracket access for properties (lib['temperature']) and for slices that return a new Library.

Files: save_to_file writes the library's state as a plain dictionary pickle (dimensions, dimension ordering,
properties, extra attributes; a mechanism among the extra attributes is stored as its mechanism data), the form the
reference's Library.load_from_file also reads (library.py:246-256 accepts dictionary pickles). load_from_file reads
such files, and libraries pickled by the reference, with a data-only reader that executes nothing from the file.
save_to_text_directory writes the reference's plain-text layout (library.py:185-244).
"""
import os
import shutil
from copy import copy, deepcopy

import numpy as np


class Dimension:
    """One independent variable of a structured library: a name (a Python identifier) and unique 1-D values."""

    def __init__(self, name, values, log_scaled=False):
        values = np.asarray(values)
        if not name.isidentifier():
            raise ValueError('Dimension name "%s" must be a valid Python identifier (no hyphens or spaces)' % name)
        if values.ndim != 1:
            raise ValueError('Dimension "%s": values must be one-dimensional' % name)
        if values.size != np.unique(values).size:
            raise ValueError('Dimension "%s": values must be unique' % name)
        self._name = name
        self._values = np.copy(values)
        self._min = np.min(values)
        self._max = np.max(values)
        self._npts = values.size
        self._log_scaled = bool(log_scaled)

    def __str__(self):
        return 'Dimension("%s" spanning [%s, %s] with %d points)' % (self._name, self._min, self._max, self._npts)

    __repr__ = __str__

    @property
    def name(self):
        return self._name

    @property
    def values(self):
        return self._values

    @property
    def min(self):
        return self._min

    @property
    def max(self):
        return self._max

    @property
    def npts(self):
        return self._npts

    @property
    def log_scaled(self):
        return self._log_scaled

    def _get_dict_for_file_save(self):
        return {'name': self._name, 'values': self._values, 'log_scaled': self._log_scaled}


class LibraryIndexError(IndexError):
    pass


class Library:
    """Properties tabulated over the grid of its dimensions (library.py:93-460)."""

    def __init__(self, *dimensions):
        dims = [d if isinstance(d, Dimension) else Dimension(d[0], d[1], False if len(d) == 2 else d[2])
                for d in dimensions]
        self._dims = {d.name: Dimension(d.name, d.values, d.log_scaled) for d in dims}
        self._dims_ordering = {i: d.name for i, d in enumerate(dims)}
        self._props = {}
        self._extra_attributes = {}
        self._grid_shape = ()
        self._grid_size = 0
        if dims:
            self._set_grid()

    def _set_grid(self):
        grid = np.meshgrid(*[self._dims[self._dims_ordering[i]].values for i in range(len(self._dims_ordering))],
                           indexing='ij')
        self._grid_shape = grid[0].shape
        self._grid_size = grid[0].size
        for i in range(len(self._dims_ordering)):
            d = self._dims[self._dims_ordering[i]]
            setattr(self, d.name + '_name', d.name)
            setattr(self, d.name + '_values', d.values)
            setattr(self, d.name + '_min', d.min)
            setattr(self, d.name + '_max', d.max)
            setattr(self, d.name + '_npts', d.npts)
            setattr(self, d.name + '_grid', np.copy(grid[i]))

    # ---------------- dimensions ----------------
    def scale_dimension(self, dim_name, multiplier):
        self.remap_dimension(dim_name, lambda x: multiplier * x)

    def remap_dimension(self, dim_name, mapping):
        if dim_name not in self._dims:
            raise KeyError('Invalid dimension name "%s"' % dim_name)
        old = self._dims[dim_name]
        self._dims[dim_name] = Dimension(dim_name, mapping(old.values), old.log_scaled)
        self._set_grid()

    @property
    def extra_attributes(self):
        return self._extra_attributes

    @property
    def size(self):
        return self._grid_size

    @property
    def shape(self):
        return self._grid_shape

    @property
    def props(self):
        return list(self._props.keys())

    @property
    def dims(self):
        return [self._dims[self._dims_ordering[i]] for i in range(len(self._dims_ordering))]

    @property
    def dim_names(self):
        return [d.name for d in self.dims]

    def dim(self, name):
        return self._dims[name]

    def get_empty_dataset(self):
        return np.ndarray(self._grid_shape)

    def add_empty_property(self, name):
        self._props[name] = self.get_empty_dataset()

    def remove(self, *quantities):
        for q in quantities:
            self._props.pop(q)

    # ---------------- data access ----------------
    def __setitem__(self, quantity, values):
        if isinstance(values, np.ndarray):
            if values.shape != self._grid_shape:
                raise ValueError('The shape of the "%s" array %s does not conform to the library grid %s'
                                 % (quantity, values.shape, self._grid_shape))
            if quantity not in self._props:
                self._props[quantity] = values.view()
            else:
                self._props[quantity][:] = values
        elif isinstance(values, (float, int)) and not isinstance(values, bool):
            if quantity not in self._props:
                self._props[quantity] = self.get_empty_dataset()
            self._props[quantity].fill(float(values))
        else:
            raise TypeError('Library values must be a numpy array or a float, received %r' % (values,))

    def __getitem__(self, *slices):
        arg = slices[0]
        if isinstance(arg, str):
            if len(slices) != 1:
                raise LibraryIndexError('Library[...] takes one property name or slices, given %s' % (slices,))
            return self._props[arg]
        if isinstance(arg, slice):
            if arg == slice(None) and len(self._dims) > 1 and len(slices) == 1:
                slices = tuple([slice(None)] * len(self._dims))
        else:
            slices = arg
        if not isinstance(slices, tuple):
            slices = (slices,)
        if len(slices) != len(self._dims):
            raise LibraryIndexError('Library[...] needs one slice per dimension (%d), given %d'
                                    % (len(self._dims), len(slices)))
        new_dims = []
        for d, s in zip(self.dims, slices):
            if not isinstance(s, (slice, int)):
                raise LibraryIndexError('Library[...] takes standard slices or integers, given %s' % (slices,))
            v = d.values[s]
            new_dims.append(Dimension(d.name, np.atleast_1d(v), d.log_scaled))
        out = Library(*new_dims)
        for p in self._props:
            out[p] = np.asarray(self._props[p][slices]).reshape(out.shape)
        for k, v in self._extra_attributes.items():
            out.extra_attributes[k] = v
        return out

    def __contains__(self, prop):
        return prop in self._props

    def __str__(self):
        lines = ['Library with %d dimensions and %d properties' % (len(self._dims), len(self._props))]
        lines += ['%d. %s' % (i + 1, d) for i, d in enumerate(self.dims)]
        lines += ['%-20s, min = %s max = %s' % (k, np.min(v), np.max(v)) for k, v in self._props.items()]
        return '\n'.join(lines)

    def __repr__(self):
        return 'Library(ndim=%d, nproperties=%d, properties=[%s])' % (len(self._dims), len(self._props),
                                                                      ', '.join(self._props))

    # ---------------- copies and reshapes ----------------
    def __copy__(self):
        out = Library(*[Dimension(d.name, d.values, d.log_scaled) for d in self.dims])
        for p in self._props:
            out[p] = self._props[p]
        out._extra_attributes = dict(self._extra_attributes)
        return out

    def __deepcopy__(self, memo=None):
        out = Library(*[Dimension(d.name, np.copy(d.values), d.log_scaled) for d in self.dims])
        for p in self._props:
            out[p] = np.copy(self._props[p])
        out._extra_attributes = {k: deepcopy(v) for k, v in self._extra_attributes.items()}
        return out

    @classmethod
    def copy(cls, library):
        return copy(library)

    @classmethod
    def deepcopy(cls, library):
        return deepcopy(library)

    @classmethod
    def squeeze(cls, library):
        """Drop dimensions with one value; with none left, a dictionary of the scalar values (library.py:293-315)."""
        keep = [d for d in library.dims if d.values.size > 1]
        if not keep:
            return {'properties': {p: np.squeeze(library[p]) for p in library.props},
                    'dimensions': {d.name: (np.squeeze(d.values), d.log_scaled) for d in library.dims},
                    'extra_attributes': library.extra_attributes}
        out = Library(*[Dimension(d.name, d.values, d.log_scaled) for d in keep])
        for p in library.props:
            out[p] = np.squeeze(library[p])
        out._extra_attributes = dict(library.extra_attributes)
        return out

    @classmethod
    def swapaxes(cls, library, idx1, idx2):
        names = library.dim_names
        if len(names) < 2:
            raise ValueError('Cannot perform swap on library with one dimension.')
        names[idx1], names[idx2] = names[idx2], names[idx1]
        out = Library(*[Dimension(n, library.dim(n).values, library.dim(n).log_scaled) for n in names])
        for p in library.props:
            out[p] = np.swapaxes(library[p], idx1, idx2)
        out._extra_attributes = dict(library.extra_attributes)
        return out

    # ---------------- files ----------------
    def __getstate__(self):
        ea = {}
        for k, v in self._extra_attributes.items():
            ea[k] = {'mech_data': v.reference_mech_data()} if hasattr(v, 'reference_mech_data') else v
        return {'dimensions': {d: self._dims[d]._get_dict_for_file_save() for d in self._dims},
                'dim_ordering': dict(self._dims_ordering), 'properties': self._props, 'extra_attributes': ea}

    def __setstate__(self, state):
        order = state['dim_ordering']
        dims = [None] * len(state['dimensions'])
        for index, name in order.items():
            d = state['dimensions'][name]
            dims[int(index)] = Dimension(d['name'], np.asarray(d['values']), d.get('log_scaled', False))
        self.__init__(*dims)
        for p, v in state['properties'].items():
            self[p] = np.asarray(v)
        for k, v in (state.get('extra_attributes') or {}).items():
            if k == 'mech_spec' and isinstance(v, dict) and 'mech_data' in v:
                from .mechanism import ChemicalMechanismSpec
                v = ChemicalMechanismSpec.from_mech_data(v['mech_data'])
            self._extra_attributes[k] = v

    def save_to_file(self, file_name):
        """The library's state as a plain dictionary pickle (numpy arrays, strings, numbers)."""
        import pickle
        with open(file_name, 'wb') as f:
            pickle.dump(self.__getstate__(), f)

    @classmethod
    def load_from_file(cls, file_name):
        """A library saved by save_to_file, or pickled by the reference's Library, read without executing anything
        from the file."""
        from .safe_pickle import Stub, load
        data = load(file_name, stub_classes=[('spitfire.chemistry.library', 'Library'),
                                             ('spitfire.chemistry.mechanism', 'ChemicalMechanismSpec')])
        if isinstance(data, Stub):
            data = data.state
        ea = data.get('extra_attributes') or {}
        for k, v in list(ea.items()):
            if isinstance(v, Stub):
                ea[k] = v.state
        out = cls()
        out.__setstate__(data)
        return out

    def save_to_text_directory(self, output_directory, ravel_order='F', format='%.14e', overwrite=False):
        """Plain-text dump in the reference's layout: metadata files naming the independent and dependent variables
        and the extra attributes, and one bulk-data file per dimension and per property (flattened in ravel_order).
        An existing directory is replaced only with overwrite=True (the reference asks interactively)."""
        if os.path.isdir(output_directory):
            if not overwrite:
                raise FileExistsError(output_directory + ' exists; pass overwrite=True to replace it')
            shutil.rmtree(output_directory)
        os.mkdir(output_directory)
        names = {p: p.replace(' ', '_') for p in self.props}
        with open(os.path.join(output_directory, 'metadata_independent_variables.txt'), 'w') as f:
            for d in self.dims:
                f.write(d.name + '\n')
        with open(os.path.join(output_directory, 'metadata_dependent_variables.txt'), 'w') as f:
            for p in self.props:
                f.write(names[p] + '\n')
        with open(os.path.join(output_directory, 'metadata_user_defined_attributes.txt'), 'w') as f:
            f.write(str(self._extra_attributes))
        for d in self.dims:
            np.savetxt(os.path.join(output_directory, 'bulkdata_ivar_%s.txt' % d.name), d.values, fmt=format)
        for p in self.props:
            np.savetxt(os.path.join(output_directory, 'bulkdata_dvar_%s.txt' % names[p]),
                       self[p].ravel(order=ravel_order), fmt=format)
