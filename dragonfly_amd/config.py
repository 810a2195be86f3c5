"""config.ini reading with the reference's file-name rules
(dragonfly/utils/py_src/read_config.py:9-24): a file name is relative to the folder of the
config file, and a value of the form 'section:::key' stands for the value of that other key.
"""
import configparser
import os.path as op

_REQUIRED = object()  # get_filename's default: a missing key is an error


class MyConfigParser(configparser.ConfigParser):
    """ConfigParser that remembers the folder of the file it read and resolves file names
    against it."""

    config_folder = ''

    def read(self, filename):
        self.config_folder = op.dirname(filename)
        return super().read(filename)

    def get_filename(self, section, option, *, fallback=_REQUIRED):
        """Path named by [section] option. Chained 'section:::key' values are followed until
        a plain name is reached, which is then joined to the config folder. A missing key
        raises NoOptionError unless a fallback is given; a fallback name is resolved like a
        value read from the file, and None or the empty string are returned as they are."""
        visited = set()
        while True:
            if (section, option) in visited:
                raise ValueError('config key %s:::%s refers back to itself' % (section, option))
            visited.add((section, option))
            if self.has_option(section, option):
                value = self.get(section, option)
            elif fallback is _REQUIRED:
                raise configparser.NoOptionError(option, section)
            else:
                value = fallback
            if not value:
                return value
            if ':::' not in value:
                return op.join(self.config_folder, value)
            section, option = value.split(':::', 1)
