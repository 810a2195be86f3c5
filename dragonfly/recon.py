"""dragonfly.recon of the reference (recon.pyx:19-445): EMCRecon and the `dragonfly.emc`
command, `python -m dragonfly.recon niter -c config.ini`."""
from dragonfly_amd.recon import EMCRecon, main

__all__ = ['EMCRecon', 'main']

if __name__ == '__main__':
    main()
