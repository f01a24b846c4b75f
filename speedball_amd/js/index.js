'use strict';
// Public entry point of the Node host (the list of js/index.js:12-42 that reaches the
// probe update): installSpeedballGI and createProbeField, over the MI355X library.
const { installSpeedballGI } = require('./install');
const { createProbeField, computeGridResolution } = require('./probe_field');
const { buildScene, uberRecord, lightRecord } = require('./scene_ingest');
const { loadAddon, STATUS, BUF, BLEND, TRAVERSAL } = require('./addon');

module.exports = {
    installSpeedballGI, createProbeField, computeGridResolution, buildScene, uberRecord, lightRecord, loadAddon,
    STATUS, BUF, BLEND, TRAVERSAL,
};
