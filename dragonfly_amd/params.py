"""EMCParams: reconstruction parameters (mirror of dragonfly/params.pyx:9-128)."""
import configparser

from .config import MyConfigParser


class EMCParams:
    """Reconstruction parameters with the reference's defaults (params.pyx:16-53) and
    config keys (params.pyx:67-128)."""

    def __init__(self):
        self.rank = 0
        self.num_proc = 1
        self.start_iter = 1
        self.known_scale = 0
        self.log_fname = 'EMC.log'
        self.output_folder = 'data/'
        self.rtype = 'RECON3D'
        self.need_scaling = 0
        self.alpha = 0.
        self.beta_factor = -1.
        self.radius = 0.
        self.num_modes = 1
        self.nonrot_modes = 0
        self.rot_per_mode = 0
        self.beta_jump = 1.
        self.beta_period = 100
        self.radius_jump = 0.
        self.radius_period = 100
        self.oversampling = 10.
        self.sigmasq = 0.
        self.friedel_sym = 0
        self.axial_sym = 1
        self.save_prob = 0
        self.update_scale = 1
        self.refine = 0
        self.coarse_div = 0
        self.fine_div = 0
        self.fixed_seed = 0
        self.verbosity = 9
        self.iteration = 1
        self.num_iter = 0

    def from_config(self, config_fname, section_name='emc'):
        config = MyConfigParser()
        config.read(config_fname)
        try:
            self.output_folder = config.get_filename(section_name, 'output_folder', fallback='data/')
        except configparser.NoOptionError:
            self.output_folder = 'data/'
        try:
            self.log_fname = config.get_filename(section_name, 'log_file', fallback='EMC.log')
        except configparser.NoOptionError:
            self.log_fname = 'EMC.log'
        rtype = config.get(section_name, 'recon_type', fallback='3d')
        if rtype not in ('3d', '2d', 'rz'):
            raise ValueError('Unknown recon_type %s' % rtype)
        if rtype != '3d':
            raise NotImplementedError('only recon_type 3d is implemented on the GPU engine')
        self.rtype = 'RECON3D'
        self.verbosity = config.getint(section_name, 'verbosity', fallback=9)
        self.need_scaling = config.getint(section_name, 'need_scaling', fallback=0)
        self.alpha = config.getfloat(section_name, 'alpha', fallback=0.)
        self.beta_factor = config.getfloat(section_name, 'beta_factor', fallback=1.)
        self.radius = config.getfloat(section_name, 'radius', fallback=0.)
        self.num_modes = config.getint(section_name, 'num_modes', fallback=1)
        self.nonrot_modes = config.getint(section_name, 'num_nonrot_modes', fallback=0)
        self.rot_per_mode = config.getint(section_name, 'num_rot', fallback=0)
        if self.num_modes < 1:
            raise ValueError('num_modes must be at least 1')
        if self.nonrot_modes != 0:
            # the Cython engine's Model holds num_modes volumes (iterate.pyx:164), so a
            # non-rotating mode (index num_modes and up, max_prob.c:82-84) has no volume
            raise NotImplementedError('num_nonrot_modes > 0 is not supported (the model has no '
                                      'volume for a non-rotating mode)')
        beta_schedule = config.get(section_name, 'beta_schedule', fallback='1. 100').split()
        self.beta_jump = float(beta_schedule[0])
        self.beta_period = int(beta_schedule[1])
        radius_schedule = config.get(section_name, 'radius_schedule', fallback='0. 100').split()
        self.radius_jump = float(radius_schedule[0])
        self.radius_period = int(radius_schedule[1])
        self.oversampling = config.getfloat(section_name, 'oversampling', fallback=10.)
        self.sigmasq = config.getfloat(section_name, 'gaussian_sigma', fallback=0.) ** 2
        self.friedel_sym = config.getint(section_name, 'friedel_sym', fallback=0)
        self.axial_sym = config.getint(section_name, 'axial_sym', fallback=1)
        if self.axial_sym > 1:
            # recon.pyx:135-137 symmetrises every mode about the axis; the engine's model
            # update applies Friedel symmetry only
            raise NotImplementedError('axial_sym > 1 is not supported by the GPU engine')
        self.save_prob = config.getint(section_name, 'save_prob', fallback=0)
        self.update_scale = config.getint(section_name, 'update_scale', fallback=1)
        num_divs = config.get(section_name, 'num_div', fallback='0').split()
        if len(num_divs) > 1:
            self.refine = 1
            self.fine_div = int(num_divs[0])
            self.coarse_div = int(num_divs[1])
        self.fixed_seed = config.getint(section_name, 'fixed_seed', fallback=0)
        return self
