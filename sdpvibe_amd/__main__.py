"""python -m sdpvibe_amd ...: the sdpb-compatible command line (cli.main)"""
import sys

from .cli import main

sys.exit(main())
