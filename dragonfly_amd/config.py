"""config.ini parsing with the reference's MyConfigParser semantics
(dragonfly/utils/py_src/read_config.py:9-24): file names are relative to the config's folder
and a value 'section:::key' chains to another key."""
import configparser
import os.path as op


class MyConfigParser(configparser.ConfigParser):
    def read(self, filename):
        self.config_folder = op.dirname(filename)
        super().read(filename)

    def get_filename(self, section, option, *, fallback='NULL'):
        fname = self.get(section, option, fallback=fallback)
        if fname == 'NULL':
            raise configparser.NoOptionError(option, section)
        if fname is None or fname == '':
            return fname
        if ':::' in fname:
            subsec, subtag = fname.split(':::')
            return self.get_filename(subsec, subtag)
        return op.join(self.config_folder, fname)
